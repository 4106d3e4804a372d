// PyTorch op boundary: torch.ops.deepmd.tabulate_fusion_se_a on HIP tensors.
//
// Same registration, schema, device contract, error messages and autograd
// chain as the reference op library
// (source/op/pt/tabulate_multi_device.cc:23-157, 764-996, 1669-1735 and
// print_summary.cc:14):
//   deepmd::tabulate_fusion_se_a(Tensor table, Tensor table_info,
//                                Tensor em_x, Tensor em, int last_layer_size) -> Tensor[]
//   forward  -> TabulateFusionSeAGradOp (backward) -> grad-grad (double backward),
//   no gradient to table / table_info, table_info read on the host.
// The kernels behind it are the dpmi_tabulate_* entry points of
// include/dpmi.h, launched on torch's current HIP stream.  The library has
// no CPU arithmetic: host tensors (the reference's CPU kernels,
// tabulate_multi_device.cc:147-152) are staged to the current HIP device,
// computed there and the results copied back, so host and device callers
// get the same HIP results; without a visible GPU the call raises.
#include <ATen/hip/HIPContext.h>
#include <torch/torch.h>

#include <initializer_list>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/dpmi.h"

namespace {

using NamedTensor = std::pair<const char*, torch::Tensor>;

void* cur_stream() { return (void*)at::hip::getCurrentHIPStream().stream(); }

void check_status(int rc) {
  if (rc == DPMI_ERR_UNSUPPORTED || rc == DPMI_ERR_ARGUMENT) TORCH_CHECK(false, dpmi_last_error());
  TORCH_CHECK(rc == DPMI_OK, "dpmi error ", rc, ": ", dpmi_last_error());
}

void check_devices(const torch::Tensor& table, const torch::Tensor& table_info,
                   std::initializer_list<NamedTensor> tensors) {
  // tabulate_multi_device.cc:61-86
  TORCH_CHECK(table_info.device().is_cpu(), "table_info must be on the CPU");
  for (const auto& [name, t] : tensors) {
    if (!t.defined()) continue;
    TORCH_CHECK(t.device() == table.device(), name, " must be on the same device as table; table is on ",
                table.device(), " but ", name, " is on ", t.device());
  }
  TORCH_CHECK(table.device().is_cuda(),
              "dpmi provides the GPU (HIP) implementation of tabulate_fusion_se_a; the table is on ",
              table.device());
}

void check_basis(int64_t nd) {
  TORCH_CHECK(nd == 4 || nd == 9 || nd == 16 || nd == 25,
              "The environment basis dimension must be 4, 9, 16, or 25, got ", nd);
}

void check_last_layer(int64_t M) {
  TORCH_CHECK(M > 0 && M <= 1024, "last_layer_size must be between 1 and 1024 for GPU tabulation, but got ", M);
}

// The reference reads every tensor through .view({-1}), which raises for a
// non-contiguous tensor; raw pointers of such a tensor would be read with the
// wrong strides, so they are rejected the same way.
template <typename T>
const T* cptr(const torch::Tensor& t) {
  TORCH_CHECK(t.is_contiguous(), "tabulate_fusion_se_a: input tensors must be contiguous");
  return t.data_ptr<T>();
}

// Host tensors run on the current HIP device (see the file comment).
torch::Tensor to_dev(const torch::Tensor& t) {
  if (!t.defined() || t.device().is_cuda()) return t;
  TORCH_CHECK(torch::cuda::is_available(),
              "dpmi computes tabulate_fusion_se_a on the GPU; host tensors need a visible HIP device");
  return t.to(torch::Device(torch::kCUDA, at::hip::current_device()));
}

template <typename T>
void fwd(const torch::Tensor& table, const torch::Tensor& info, const torch::Tensor& em_x, const torch::Tensor& em,
         int64_t M, torch::Tensor& out) {
  check_devices(table, info, {{"em_x", em_x}, {"em", em}, {"descriptor", out}});
  if (table.dim() != 2) throw std::invalid_argument("Dim of table should be 2");
  if (em_x.dim() != 2) throw std::invalid_argument("Dim of input should be 2");
  if (em.dim() != 3) throw std::invalid_argument("Dim of input should be 3");
  const int64_t nloc = em.size(0), nnei = em.size(1), nd = em.size(2);
  check_basis(nd);
  check_last_layer(M);
  auto info_c = info.contiguous();
  int rc;
  if constexpr (std::is_same_v<T, double>)
    rc = dpmi_tabulate_fusion_se_a_f64(out.data_ptr<double>(), cptr<double>(table), cptr<double>(info_c),
                                       cptr<double>(em_x), cptr<double>(em), nloc, nnei, M, 1, nd, cur_stream());
  else
    rc = dpmi_tabulate_fusion_se_a_f32(out.data_ptr<float>(), cptr<float>(table), cptr<float>(info_c),
                                       cptr<float>(em_x), cptr<float>(em), nloc, nnei, M, 1, nd, cur_stream());
  check_status(rc);
}

template <typename T>
void grad(const torch::Tensor& table, const torch::Tensor& info, const torch::Tensor& em_x, const torch::Tensor& em,
          const torch::Tensor& dy, const torch::Tensor& descriptor, torch::Tensor& dy_dem_x, torch::Tensor& dy_dem) {
  check_devices(table, info, {{"em_x", em_x}, {"em", em}, {"dy", dy}, {"descriptor", descriptor},
                              {"dy_dem_x", dy_dem_x}, {"dy_dem", dy_dem}});
  if (dy.dim() != 3) throw std::invalid_argument("Dim of dy_tensor should be 3");
  const int64_t nloc = em.size(0), nnei = em.size(1), nd = em.size(2), M = descriptor.size(2);
  check_basis(nd);
  check_last_layer(M);
  auto info_c = info.contiguous();
  int rc;
  if constexpr (std::is_same_v<T, double>)
    rc = dpmi_tabulate_fusion_se_a_grad_f64(dy_dem_x.data_ptr<double>(), dy_dem.data_ptr<double>(),
                                            cptr<double>(table), cptr<double>(info_c), cptr<double>(em_x),
                                            cptr<double>(em), cptr<double>(dy), nloc, nnei, M, 1, nd, cur_stream());
  else
    rc = dpmi_tabulate_fusion_se_a_grad_f32(dy_dem_x.data_ptr<float>(), dy_dem.data_ptr<float>(),
                                            cptr<float>(table), cptr<float>(info_c), cptr<float>(em_x),
                                            cptr<float>(em), cptr<float>(dy), nloc, nnei, M, 1, nd, cur_stream());
  check_status(rc);
}

template <typename T>
void grad_grad(const torch::Tensor& table, const torch::Tensor& info, const torch::Tensor& em_x,
               const torch::Tensor& em, const torch::Tensor& dz_dx, const torch::Tensor& dz_dem,
               const torch::Tensor& descriptor, bool is_sorted, torch::Tensor& dz_dy) {
  check_devices(table, info, {{"em_x", em_x}, {"em", em}, {"dz_dy_dem_x", dz_dx}, {"dz_dy_dem", dz_dem},
                              {"descriptor", descriptor}, {"dz_dy", dz_dy}});
  if (dz_dx.dim() != 2) throw std::invalid_argument("Dim of dz_dy_dem_x should be 2");
  if (dz_dem.dim() != 3) throw std::invalid_argument("Dim of dz_dy_dem should be 3");
  const int64_t nloc = em.size(0), nnei = em.size(1), nd = em.size(2), M = descriptor.size(2);
  check_basis(nd);
  check_last_layer(M);
  auto info_c = info.contiguous();
  int rc;
  if constexpr (std::is_same_v<T, double>)
    rc = dpmi_tabulate_fusion_se_a_grad_grad_f64(dz_dy.data_ptr<double>(), cptr<double>(table), cptr<double>(info_c),
                                                 cptr<double>(em_x), cptr<double>(em), cptr<double>(dz_dx),
                                                 cptr<double>(dz_dem), nloc, nnei, M, is_sorted, nd, cur_stream());
  else
    rc = dpmi_tabulate_fusion_se_a_grad_grad_f32(dz_dy.data_ptr<float>(), cptr<float>(table), cptr<float>(info_c),
                                                 cptr<float>(em_x), cptr<float>(em), cptr<float>(dz_dx),
                                                 cptr<float>(dz_dem), nloc, nnei, M, is_sorted, nd, cur_stream());
  check_status(rc);
}

class GradOp : public torch::autograd::Function<GradOp> {
 public:
  static torch::autograd::variable_list forward(torch::autograd::AutogradContext* ctx, const torch::Tensor& table,
                                                const torch::Tensor& info, const torch::Tensor& em_x,
                                                const torch::Tensor& em, const torch::Tensor& dy,
                                                const torch::Tensor& descriptor) {
    const bool host = table.device().is_cpu();
    const torch::Tensor tb = to_dev(table), ex = to_dev(em_x), e = to_dev(em), d = to_dev(dy),
                        de = to_dev(descriptor);
    auto dy_dem_x = torch::zeros_like(ex);
    auto dy_dem = torch::zeros_like(e);
    if (tb.dtype() == torch::kDouble)
      grad<double>(tb, info, ex, e, d, de, dy_dem_x, dy_dem);
    else
      grad<float>(tb, info, ex, e, d, de, dy_dem_x, dy_dem);
    ctx->save_for_backward({table, info, em_x, em, descriptor});
    if (host) return {dy_dem_x.cpu(), dy_dem.cpu()};
    return {dy_dem_x, dy_dem};
  }
  static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::variable_list g) {
    auto s = ctx->get_saved_variables();
    const bool host = s[0].device().is_cpu();
    for (size_t k = 0; k < s.size(); ++k)
      if (k != 1) s[k] = to_dev(s[k]);
    auto dz_dx = g[0].defined() ? to_dev(g[0].contiguous()) : torch::zeros_like(s[2]);
    auto dz_dem = g[1].defined() ? to_dev(g[1].contiguous()) : torch::zeros_like(s[3]);
    auto dz_dy = torch::empty_like(s[4]);
    if (s[0].dtype() == torch::kDouble)
      grad_grad<double>(s[0], s[1], s[2], s[3], dz_dx, dz_dem, s[4], true, dz_dy);
    else
      grad_grad<float>(s[0], s[1], s[2], s[3], dz_dx, dz_dem, s[4], true, dz_dy);
    return {at::Tensor(), at::Tensor(), at::Tensor(), at::Tensor(), host ? dz_dy.cpu() : dz_dy, at::Tensor()};
  }
};

class FwdOp : public torch::autograd::Function<FwdOp> {
 public:
  static torch::autograd::variable_list forward(torch::autograd::AutogradContext* ctx, const torch::Tensor& table,
                                                const torch::Tensor& info, const torch::Tensor& em_x,
                                                const torch::Tensor& em, int64_t M) {
    check_last_layer(M);
    const bool host = table.device().is_cpu() && em_x.device().is_cpu() && em.device().is_cpu();
    const torch::Tensor tb = host ? to_dev(table) : table, ex = host ? to_dev(em_x) : em_x,
                        e = host ? to_dev(em) : em;
    auto opts = torch::TensorOptions().dtype(tb.dtype()).device(tb.device());
    auto out = torch::empty({e.size(0), e.size(2), M}, opts);
    if (tb.dtype() == torch::kDouble)
      fwd<double>(tb, info, ex, e, M, out);
    else
      fwd<float>(tb, info, ex, e, M, out);
    ctx->save_for_backward({table, info, em_x, em, out});
    return {host ? out.cpu() : out};
  }
  static torch::autograd::variable_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::variable_list g) {
    auto s = ctx->get_saved_variables();
    auto dy = g[0].contiguous();
    auto r = GradOp::apply(s[0], s[1], s[2], s[3], dy, s[4]);
    return {at::Tensor(), at::Tensor(), r[0], r[1], at::Tensor()};
  }
};

std::vector<torch::Tensor> tabulate_fusion_se_a(const torch::Tensor& table, const torch::Tensor& table_info,
                                                const torch::Tensor& em_x, const torch::Tensor& em,
                                                int64_t last_layer_size) {
  return FwdOp::apply(table, table_info, em_x, em, last_layer_size);
}

torch::Tensor enable_mpi() { return torch::zeros({1}, torch::kBool); }

}  // namespace

TORCH_LIBRARY(deepmd, m) {
  m.def("enable_mpi", enable_mpi);
  m.def("tabulate_fusion_se_a", tabulate_fusion_se_a);
}
