// Host-side model and engine objects behind dpmi_sea_* (include/dpmi.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dpmi.h"
#include "geometry.h"
#include "kernels.h"

namespace dpmi {

std::vector<int4> fitting_tile_order(const std::vector<int>& tstart, int ntypes, int rows_per_tile, int nx,
                                     int cus);

struct SeaModel {
  explicit SeaModel(const dpmi_sea_model_desc& d);
  void build_mixed(const dpmi_sea_model_desc& d);
  ~SeaModel();
  int ntypes = 0, nnei = 0, M = 0, A = 0, nnets = 0, nlayers = 0;
  float rcut = 0, rcut_smth = 0;
  bool type_one_side = false;
  std::vector<int> sel, sec, net_of, width, act, res_mode, has_idt;
  std::vector<double> davg, dstd, info_host, bias_atom_e, out_b_host;
  std::vector<double> table_host;  // all nets, rows [M][6] as given
  std::vector<std::vector<double>> out_w_host;
  std::vector<long long> tab_off;
  // device
  double* d_table = nullptr;
  long long* d_tab_off = nullptr;
  char* d_infos = nullptr;
  int* d_net_of = nullptr;
  int* d_sec = nullptr;
  double* d_davg = nullptr;
  double* d_dstd = nullptr;
  std::vector<double*> fit_w, fit_b, fit_idt;  // [ntypes * nlayers]; fit_w as FP64 MFMA B fragments
  std::vector<double*> fit_wt;                 // W^T ([n_out][n_in]) as B fragments, for the backward
  // mixed-precision engine: FP32 tables, FP16 fitting weights padded to
  // multiples of 64, FP64 energy shift per type (output bias + bias_atom_e)
  std::string mixed_unsupported;  // empty when DPMI_MIXED can run this model
  std::vector<int> kpad;          // [nlayers + 1] padded widths
  float* d_table_f = nullptr;
  float* d_table_pk = nullptr;  // packed records (a0, a1 FP32; a2..a5 FP16) per column
  long long* d_tab_off_pk = nullptr;
  char* d_infos_f = nullptr;
  double* d_shift = nullptr;
  std::vector<Fit16Layer> f16;  // [ntypes * (nlayers - 1)] hidden layers
  std::vector<void*> f16_mem;
  bool fused = false;             // the fused fitting kernel covers this net
  FusedLayer* d_fused = nullptr;  // [ntypes * (nlayers - 1)]
  // matrix-core embedding: split FP16 tables (forward and adjoint layouts)
  std::vector<int> nspline;  // table rows per net
  bool embed_mfma = false;
  _Float16* d_tab16 = nullptr;
  long long* d_tab16_off = nullptr;
  float* d_tab16_pscale = nullptr;
  // row-run adjoint on the matrix cores: run tables of every net
  bool embed_runs = false;
  _Float16* d_rtab = nullptr;
  long long* d_rtab_off = nullptr;
  float* d_rtab_pscale = nullptr;
  _Float16* d_ftab = nullptr;  // forward run tables
  long long* d_ftab_off = nullptr;
  float* d_ftab_scale = nullptr;
  std::vector<std::string> type_map;  // names of the types (model files)
  // uncompressed embedding (FP64 engines): per (net, layer) the weights as
  // FP64 MFMA B fragments (W and W^T), bias and timestep
  bool compressed = true;
  int enl = 0;
  std::vector<int> ewidth, eact, eres;
  std::vector<double*> ew, ewt, eb, eidt;  // [nnets * enl]
  double** d_eidt_top = nullptr;           // [nnets] timestep of each net's last layer (or null)
};

struct SeaEngine {
  SeaEngine(const SeaModel* m, int nframes, const int* natoms, const int* atype, int precision,
            double nall_per_atom);
  ~SeaEngine();
  // One E+F+V step.  With growth on (the default) an eager step reads its
  // status word back and, when the ghost or neighbour-row capacity
  // overflowed, grows that capacity and recomputes the step.
  void compute(const double* coord, const double* box_h, double* energy, double* force, double* virial,
               double* atom_energy, double* atom_virial, hipStream_t st);
  // device status bits (DeviceFlag) of the steps since the last call; cleared
  int status_flags();
  // forward_lower: one frame whose nloc locals (the engine's atoms, same
  // types) come first among nall caller-owned extended atoms; candidate rows
  // (any order, any radius >= rcut) at jlist + r * jstride; outputs on the
  // extended atoms (no ghost fold)
  void compute_lower(const double* coord_ext, const int* atype_ext, int nall, int inum, const int* ilist,
                     const int* numneigh, const int* jlist, int jstride, double* energy, double* force_ext,
                     double* virial, double* atom_energy, double* atom_virial_ext, hipStream_t st);
  bool grow = true;
  int regrows = 0;  // capacity increases made by compute()
  long long fetch(int frame, int what, void* dst);
  // one step (eager) then the same step captured into a hipGraph; replays
  // read the coordinates at `coord` and use the boxes given here
  void capture(const double* coord, const double* box_h, double* energy, double* force, double* virial,
               double* atom_energy, hipStream_t st);
  void launch_graph(hipStream_t st);
  // the same with the boxes in device memory [nframes, 9]: every replay
  // re-reads them (NPT), provided the cell grid stays the one of the capture
  void capture_device_box(const double* coord, const double* box_dev, double* energy, double* force, double* virial,
                          double* atom_energy, hipStream_t st);
  // picks the launch-shape knobs for this engine's size bucket (and caches
  // them for later engines of the same bucket)
  void autotune(const double* coord, const double* box_h, int reps, hipStream_t st);
  Tune tune;
  int timing_report(double* ms, int max_stages, char* names, int names_len);

  const SeaModel* model;
  int nframes, precision;
  double nall_per_atom;
  int ntot = 0;
  std::vector<int> natoms, loc_off, atype, tstart;
  std::string stage_names;
  bool timing = false;

 private:
  void prepare_geometry(const double* box_h);
  void load_cached_tune();
  void mark(hipStream_t st, int stage);
  void enqueue_step(const double* coord, const double* box_h, double* energy, double* force, double* virial,
                    double* atom_energy, double* atom_virial, hipStream_t st);
  // a captured graph refers to the buffers and boxes of its capture: any
  // reallocation or new box drops it (launch_graph then asks for a capture)
  void drop_graph();
  bool capturing = false;
  const double* graph_box = nullptr;  // set while capturing for device boxes
  bool graph_dbox = false;            // the captured graph re-reads its boxes from device memory
  int sticky = 0;               // status bits read back by growing steps, reported by status()
  int* stat_pinned = nullptr;   // [1 + nframes]: status word and per-frame nall of the last growing step
  std::vector<int> ext_need_h;  // per-frame extended capacity learned from overflows

  std::vector<FrameGeom> geom_h;
  FrameGeom* geom_pinned = nullptr;
  std::vector<int> cell_frame_h, ext_cap_h;
  int ncells_tot = 0, nx_tot = 0;
  long long next_tot = 0;
  int nlist_cap = 256;
  hipStream_t last_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<hipEvent_t> events;
  int nstages = 0;
  int timed_calls = 0;
  bool geom_uploaded = false;

 public:
  void reset_timing() { timed_calls = 0; }

 private:

  FrameGeom* d_geom = nullptr;
  int *d_frame_of = nullptr, *d_center = nullptr, *d_type = nullptr, *d_pos = nullptr, *d_perm = nullptr;
  int *d_flag = nullptr, *d_nall = nullptr, *d_frame_ext_off = nullptr;
  int *d_nlist = nullptr, *d_ncand = nullptr, *d_cell_of = nullptr, *d_sorted_loc = nullptr;
  int *d_cell_frame = nullptr, *d_xcell_frame = nullptr, *d_cell_cnt = nullptr, *d_cell_start = nullptr;
  int *d_cell_slot = nullptr, *d_cell_list = nullptr;
  bool counts_clean = false;  // d_cell_cnt is zero (left so by the last complete build)
  int *d_xcnt = nullptr, *d_xstart = nullptr, *d_xgcnt = nullptr, *d_xgstart = nullptr;
  int* d_cell_sub = nullptr;  // [nx_tot, kAxisBins] record table of the cells' sub-cells
  int *d_type_ext = nullptr, *d_map_ext = nullptr, *d_xframe_of_ext = nullptr;
  CellAtom* d_cell_atoms = nullptr;
  double *d_coord_ext = nullptr, *d_force_ext = nullptr;
  double *d_em = nullptr, *d_em_deriv = nullptr, *d_rij = nullptr, *d_net_deriv = nullptr;
  bool recompute_slots = false;  // em_deriv / rij rebuilt by the force kernel (not stored)
  double *d_xyz = nullptr, *d_D = nullptr, *d_gD = nullptr, *d_atom_e = nullptr,
         *d_vpart = nullptr;
  std::vector<double*> d_fy, d_fx, d_outw;
  double *d_g0 = nullptr, *d_g1 = nullptr;
  double* d_gz[2] = {nullptr, nullptr};  // FP64 fitting backward: dE/dz ping-pong
  // mixed engine
  float *d_em_f = nullptr, *d_em_deriv_f = nullptr, *d_rij_f = nullptr, *d_net_deriv_f = nullptr;
  float *d_xyz_f = nullptr, *d_gD_f = nullptr, *d_eperm = nullptr, *d_ga = nullptr, *d_gb = nullptr;
  _Float16* d_Dh = nullptr;
  // layer 0 of the fitting in its own kernels: x_1, idt act'(z_0), dE/dz_0
  // [ntot, kpad[1]] FP16 (models with at least two hidden layers)
  _Float16 *d_fx1 = nullptr, *d_fdv1 = nullptr, *d_fgz0 = nullptr;
  int nxcd = 1;             // XCDs of the device (workgroups dealt round-robin)
  int4* d_tiles = nullptr;  // fused fitting tiles of 64 rows
  int ntiles = 0;
  int4* d_tiles32 = nullptr;  // and of 32 rows
  int ntiles32 = 0;
  int fit_rows_default = 64;
  int4* d_etiles = nullptr;  // embedding tiles: 4 descriptor positions of one type
  int netiles = 0;
  std::vector<float*> d_y16;
  std::vector<_Float16*> d_x16;
  void compute_mixed_descriptor_fitting(hipStream_t st, int& stage);
  EnvArgs env_args() const;
  EmbedArgs embed_args() const;
  void compute_fp64_descriptor_fitting(hipStream_t st, int& stage);
  int nchunks = 0;
  int *d_chunks = nullptr, *d_chunk_start = nullptr, *d_ticket = nullptr;
  double* d_part = nullptr;
  // uncompressed embedding: atoms in chunks of descriptor positions whose
  // neighbour slots fit the row buffers; rows laid out net by net
  struct NetChunk {
    int q0 = 0, q1 = 0;                 // descriptor positions [q0, q1)
    std::vector<long long> p0, rows;    // per net
  };
  std::vector<NetChunk> net_chunks;
  long long net_rows_max = 0;
  int* d_rowbase = nullptr;  // [ntot * ntypes] row of (atom, block) in its chunk (-1: none)
  double *d_nx0 = nullptr, *d_nds = nullptr, *d_ngg = nullptr;
  double* d_ndz[2] = {nullptr, nullptr};
  double* d_ngx[2] = {nullptr, nullptr};
  std::vector<double*> d_ny, d_nxl;  // per embedding layer: act output, layer output
  void net_setup();
  void net_forward(const NetChunk& c, hipStream_t st);
  void net_descriptor(hipStream_t st);
  void net_adjoint(hipStream_t st);

 public:
  // upper bound of the uncompressed embedding's row buffers (bytes); takes
  // effect for engines created afterwards in this process
  static long long net_chunk_bytes;
};

}  // namespace dpmi

struct dpmi_sea_model {
  dpmi::SeaModel* impl;
};
struct dpmi_sea_engine {
  dpmi::SeaEngine* impl;
};
