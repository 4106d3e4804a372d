// deepmd:: GPU function templates of the se_e2_a path, signature-identical to
// the reference library headers, implemented by libdpmi.so on gfx950.
//
// Reference declarations replaced (paths relative to the reference checkout):
//   struct InputNlist                    source/lib/include/neighbor_list.h:20-119
//   struct Region, init_region_cpu       source/lib/include/region.h:6-21
//   deepmd_exception{,_oom,_nlist_capacity} source/lib/include/errors.h:11-31
//   normalize_coord_gpu                  source/lib/include/coord.h:50-54
//   copy_coord_gpu                       source/lib/include/coord.h:56-82
//   build_nlist_gpu                      source/lib/include/neighbor_list.h:230-247
//   prod_env_mat_a_gpu                   source/lib/include/prod_env_mat.h:91-133
//   tabulate_fusion_se_a{,_grad,_grad_grad}_gpu  source/lib/include/tabulate.h:175-218
//   prod_force_a_gpu                     source/lib/include/prod_force.h:65-73
//   prod_virial_a_gpu                    source/lib/include/prod_virial.h:30-39
//   prod_force_grad_a_gpu                source/lib/include/prod_force_grad.h:27-34
//   prod_virial_grad_a_gpu               source/lib/include/prod_virial_grad.h:27-34
//   compute_cell_info                    source/lib/include/coord.h:39-47
//   format_nbor_list_gpu                 source/lib/include/fmt_nlist.h:22-50
//   validate_nlist_gpu_capacity, convert_nlist_gpu_device,
//   free_nlist_gpu_device, use_nlist_map source/lib/include/neighbor_list.h:187-232
//   env_mat_nbor_update                  source/lib/include/prod_env_mat.h:179-184
// Instantiated for double and float.  The float instantiations of
// copy_coord_gpu and prod_env_mat_a_gpu evaluate the geometry in FP64 from the
// FP32 inputs and round each output once; normalize_coord_gpu<float>,
// build_nlist_gpu<float> and the neighbour formatting compute in FP32 as the
// reference does.
// Contracts kept: device pointers for arrays and for Region::boxt/rec_boxt,
// host pointer for table_info, synchronous completion on the default stream,
// outputs zero-filled by the callee, capacity failures reported by return
// value 1 (copy_coord_gpu, build_nlist_gpu) or deepmd_exception_nlist_capacity,
// HIP failures by deepmd_exception.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace deepmd {

struct deepmd_exception : public std::runtime_error {
 public:
  deepmd_exception() : runtime_error("DeePMD-kit Error!") {}
  deepmd_exception(const std::string& msg) : runtime_error(std::string("DeePMD-kit Error: ") + msg) {}
};
struct deepmd_exception_oom : public deepmd_exception {
 public:
  deepmd_exception_oom() : deepmd_exception("DeePMD-kit OOM!") {}
  deepmd_exception_oom(const std::string& msg) : deepmd_exception(std::string("DeePMD-kit OOM: ") + msg) {}
};
struct deepmd_exception_nlist_capacity : public deepmd_exception {
 public:
  using deepmd_exception::deepmd_exception;
};

struct InputNlist {
  int inum;
  int* ilist;
  int* numneigh;
  int** firstneigh;
  int nswap;
  int* sendnum;
  int* recvnum;
  int* firstrecv;
  int** sendlist;
  int* sendproc;
  int* recvproc;
  void* world;
  int mask = 0xFFFFFFFF;
  int* mapping = nullptr;
  int nprocs = 1;
  InputNlist()
      : inum(0), ilist(nullptr), numneigh(nullptr), firstneigh(nullptr), nswap(0), sendnum(nullptr),
        recvnum(nullptr), firstrecv(nullptr), sendlist(nullptr), sendproc(nullptr), recvproc(nullptr),
        world(nullptr) {}
  InputNlist(int inum_, int* ilist_, int* numneigh_, int** firstneigh_)
      : inum(inum_), ilist(ilist_), numneigh(numneigh_), firstneigh(firstneigh_), nswap(0), sendnum(nullptr),
        recvnum(nullptr), firstrecv(nullptr), sendlist(nullptr), sendproc(nullptr), recvproc(nullptr),
        world(nullptr) {}
  void set_mask(int mask_) { mask = mask_; }
  void set_mapping(int* mapping_) { mapping = mapping_; }
};

template <typename FPTYPE>
struct Region {
  FPTYPE* boxt;
  FPTYPE* rec_boxt;
  Region();
  Region(FPTYPE* extern_boxt, FPTYPE* extern_rec_boxt);
  ~Region();

 private:
  bool self_allocated;
};

template <typename FPTYPE>
void init_region_cpu(Region<FPTYPE>& region, const FPTYPE* boxt);

template <typename FPTYPE>
void normalize_coord_gpu(FPTYPE* coord, const int natom, const Region<FPTYPE>& region);

template <typename FPTYPE>
int copy_coord_gpu(FPTYPE* out_c, int* out_t, int* mapping, int* nall, int* int_data, const FPTYPE* in_c,
                   const int* in_t, const int& nloc, const int& mem_nall, const int& loc_cellnum,
                   const int& total_cellnum, const int* cell_info, const Region<FPTYPE>& region);

template <typename FPTYPE>
int build_nlist_gpu(InputNlist& nlist, int* max_list_size, int* nlist_data, const FPTYPE* c_cpy, const int& nloc,
                    const int& nall, const int& mem_size, const float& rcut, const int& nframes = 1,
                    const int* type = nullptr);

// host Region (the reference computes the grid on the host); cell_info[23]:
// nat_stt, ncell, ext_stt, ext_end, ngcell, cell_shift, cell_iter,
// loc_cellnum, total_cellnum
template <typename FPTYPE>
void compute_cell_info(int* cell_info, const float& rcut, const Region<FPTYPE>& region);

int validate_nlist_gpu_capacity(const InputNlist& nlist, const int row_capacity);
void convert_nlist_gpu_device(InputNlist& gpu_nlist, InputNlist& cpu_nlist, int*& gpu_memory,
                              const int& max_nbor_size);
void free_nlist_gpu_device(InputNlist& gpu_nlist);
void use_nlist_map(int* nlist, const int* nlist_map, const int nloc, const int nnei);
void env_mat_nbor_update(InputNlist& inlist, InputNlist& gpu_inlist, int& max_nbor_size, int*& nbor_list_dev,
                         const int* mesh, const int size);

template <typename FPTYPE>
void format_nbor_list_gpu(int* nlist, const FPTYPE* coord, const int* type, const InputNlist& gpu_inlist,
                          int* array_int, unsigned long long* array_longlong, const int max_nbor_size, const int nloc,
                          const int nall, const int nframes, const float rcut, const std::vector<int> sec);
template <typename FPTYPE>
inline void format_nbor_list_gpu(int* nlist, const FPTYPE* coord, const int* type, const InputNlist& gpu_inlist,
                                 int* array_int, unsigned long long* array_longlong, const int max_nbor_size,
                                 const int nloc, const int nall, const float rcut, const std::vector<int> sec) {
  format_nbor_list_gpu(nlist, coord, type, gpu_inlist, array_int, array_longlong, max_nbor_size, nloc, nall, 1, rcut,
                       sec);
}

template <typename FPTYPE>
void prod_env_mat_a_gpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord, const int* type,
                        const InputNlist& gpu_inlist, int* array_int, unsigned long long* array_longlong,
                        const int max_nbor_size, const FPTYPE* avg, const FPTYPE* std, const int nloc,
                        const int nall, const int nframes, const float rcut, const float rcut_smth,
                        const std::vector<int> sec, const int* f_type = nullptr);

template <typename FPTYPE>
inline void prod_env_mat_a_gpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord,
                               const int* type, const InputNlist& gpu_inlist, int* array_int,
                               unsigned long long* array_longlong, const int max_nbor_size, const FPTYPE* avg,
                               const FPTYPE* std, const int nloc, const int nall, const float rcut,
                               const float rcut_smth, const std::vector<int> sec, const int* f_type = nullptr) {
  prod_env_mat_a_gpu(em, em_deriv, rij, nlist, coord, type, gpu_inlist, array_int, array_longlong, max_nbor_size,
                     avg, std, nloc, nall, 1, rcut, rcut_smth, sec, f_type);
}

template <typename FPTYPE>
void tabulate_fusion_se_a_gpu(FPTYPE* out, const FPTYPE* table, const FPTYPE* table_info, const FPTYPE* em_x,
                              const FPTYPE* em, const FPTYPE* two_embed, const int nloc, const int nnei,
                              const int last_layer_size, const bool is_sorted = true, const int ndescrpt = 4);

template <typename FPTYPE>
void tabulate_fusion_se_a_grad_gpu(FPTYPE* dy_dem_x, FPTYPE* dy_dem, FPTYPE* dy_dtwo, const FPTYPE* table,
                                   const FPTYPE* table_info, const FPTYPE* em_x, const FPTYPE* em,
                                   const FPTYPE* two_embed, const FPTYPE* dy, const int nloc, const int nnei,
                                   const int last_layer_size, const bool is_sorted = true, const int ndescrpt = 4);

template <typename FPTYPE>
void tabulate_fusion_se_a_grad_grad_gpu(FPTYPE* dz_dy, const FPTYPE* table, const FPTYPE* table_info,
                                        const FPTYPE* em_x, const FPTYPE* em, const FPTYPE* two_embed,
                                        const FPTYPE* dz_dy_dem_x, const FPTYPE* dz_dy_dem,
                                        const FPTYPE* dz_dy_dtwo, const int nloc, const int nnei,
                                        const int last_layer_size, const bool is_sorted = true,
                                        const int ndescrpt = 4);

template <typename FPTYPE>
void prod_force_a_gpu(FPTYPE* force, const FPTYPE* net_deriv, const FPTYPE* in_deriv, const int* nlist,
                      const int nloc, const int nall, const int nnei, const int nframes);

template <typename FPTYPE>
void prod_virial_a_gpu(FPTYPE* virial, FPTYPE* atom_virial, const FPTYPE* net_deriv, const FPTYPE* env_deriv,
                       const FPTYPE* rij, const int* nlist, const int nloc, const int nall, const int nnei);

template <typename FPTYPE>
void prod_force_grad_a_gpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const int* nlist,
                           const int nloc, const int nnei, const int nframes);

template <typename FPTYPE>
void prod_virial_grad_a_gpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const FPTYPE* rij,
                            const int* nlist, const int nloc, const int nnei);

// ---------------------------------------------------------------------------
// Host-array entry points (the reference's *_cpu functions).  This library has
// no CPU arithmetic: each call stages its host arrays to the current HIP
// device, runs the *_gpu path above and copies the results back.
//   normalize_coord_cpu, copy_coord_cpu          source/lib/include/coord.h:9-37
//   build_nlist_cpu                              source/lib/include/neighbor_list.h:166-176
//   format_nlist_cpu                             source/lib/include/fmt_nlist.h:11-19
//   prod_env_mat_a_cpu                           source/lib/include/prod_env_mat.h:10-49
//   tabulate_fusion_se_a{,_grad,_grad_grad}_cpu  source/lib/include/tabulate.h:28-71
//   prod_force_a_cpu                             source/lib/include/prod_force.h:20-27
//   prod_virial_a_cpu                            source/lib/include/prod_virial.h:7-15
//   prod_force_grad_a_cpu                        source/lib/include/prod_force_grad.h:7-13
//   prod_virial_grad_a_cpu                       source/lib/include/prod_virial_grad.h:7-13
template <typename FPTYPE>
void normalize_coord_cpu(FPTYPE* coord, const int natom, const Region<FPTYPE>& region);
template <typename FPTYPE>
int copy_coord_cpu(FPTYPE* out_c, int* out_t, int* mapping, int* nall, const FPTYPE* in_c, const int* in_t,
                   const int& nloc, const int& mem_nall, const float& rcut, const Region<FPTYPE>& region);
template <typename FPTYPE>
int build_nlist_cpu(InputNlist& nlist, int* max_list_size, const FPTYPE* c_cpy, const int& nloc, const int& nall,
                    const int& mem_size, const float& rcut, const int& nframes = 1, const int* type = nullptr);
template <typename FPTYPE>
void format_nlist_cpu(int* nlist, const InputNlist& in_nlist, const FPTYPE* coord, const int* type, const int nloc,
                      const int nall, const float rcut, const std::vector<int> sec);
template <typename FPTYPE>
void prod_env_mat_a_cpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord, const int* type,
                        const InputNlist& inlist, const int max_nbor_size, const FPTYPE* avg, const FPTYPE* std,
                        const int nloc, const int nall, const int nframes, const float rcut, const float rcut_smth,
                        const std::vector<int> sec, const int* f_type = nullptr);
template <typename FPTYPE>
inline void prod_env_mat_a_cpu(FPTYPE* em, FPTYPE* em_deriv, FPTYPE* rij, int* nlist, const FPTYPE* coord,
                               const int* type, const InputNlist& inlist, const int max_nbor_size, const FPTYPE* avg,
                               const FPTYPE* std, const int nloc, const int nall, const float rcut,
                               const float rcut_smth, const std::vector<int> sec, const int* f_type = nullptr) {
  prod_env_mat_a_cpu(em, em_deriv, rij, nlist, coord, type, inlist, max_nbor_size, avg, std, nloc, nall, 1, rcut,
                     rcut_smth, sec, f_type);
}
template <typename FPTYPE>
void tabulate_fusion_se_a_cpu(FPTYPE* out, const FPTYPE* table, const FPTYPE* table_info, const FPTYPE* em_x,
                              const FPTYPE* em, const FPTYPE* two_embed, const int nloc, const int nnei,
                              const int last_layer_size, const bool is_sorted = true, const int ndescrpt = 4);
template <typename FPTYPE>
void tabulate_fusion_se_a_grad_cpu(FPTYPE* dy_dem_x, FPTYPE* dy_dem, FPTYPE* dy_dtwo, const FPTYPE* table,
                                   const FPTYPE* table_info, const FPTYPE* em_x, const FPTYPE* em,
                                   const FPTYPE* two_embed, const FPTYPE* dy, const int nloc, const int nnei,
                                   const int last_layer_size, const bool is_sorted = true, const int ndescrpt = 4);
template <typename FPTYPE>
void tabulate_fusion_se_a_grad_grad_cpu(FPTYPE* dz_dy, const FPTYPE* table, const FPTYPE* table_info,
                                        const FPTYPE* em_x, const FPTYPE* em, const FPTYPE* two_embed,
                                        const FPTYPE* dz_dy_dem_x, const FPTYPE* dz_dy_dem, const FPTYPE* dz_dy_dtwo,
                                        const int nloc, const int nnei, const int last_layer_size,
                                        const bool is_sorted = true, const int ndescrpt = 4);
template <typename FPTYPE>
void prod_force_a_cpu(FPTYPE* force, const FPTYPE* net_deriv, const FPTYPE* in_deriv, const int* nlist,
                      const int nloc, const int nall, const int nnei, const int nframes);
template <typename FPTYPE>
void prod_virial_a_cpu(FPTYPE* virial, FPTYPE* atom_virial, const FPTYPE* net_deriv, const FPTYPE* env_deriv,
                       const FPTYPE* rij, const int* nlist, const int nloc, const int nall, const int nnei);
template <typename FPTYPE>
void prod_force_grad_a_cpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const int* nlist,
                           const int nloc, const int nnei, const int nframes);
template <typename FPTYPE>
void prod_virial_grad_a_cpu(FPTYPE* grad_net, const FPTYPE* grad, const FPTYPE* env_deriv, const FPTYPE* rij,
                            const int* nlist, const int nloc, const int nnei);

}  // namespace deepmd
