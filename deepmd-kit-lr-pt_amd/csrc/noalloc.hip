// Fixed device scratch for the allocation-free entry point
// dpmi_prod_virial_a_f64_noalloc (capi.cpp).
//
// The per-row virial partials of prod_virial_a ([nloc, 9] doubles, summed by
// a second kernel) live in a file-scope device array instead of a buffer the
// library allocates.  Each GPU holds its own copy of the array, loaded with
// the library's code object, until the process ends.  In host code the
// array's name is only a host-side stand-in, so its device address is looked
// up once per device with hipGetSymbolAddress and that address is what the
// kernels receive.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "dpmi_common.h"
#include "kernels.h"

namespace dpmi {

namespace {

static __device__ __align__(256) double scratch[9 * (size_t)kNoAllocVirialRows];

std::mutex g_mu;
std::vector<double*>& device_ptrs() {
  static auto* v = new std::vector<double*>();
  return *v;
}

}  // namespace

double* noalloc_virial_scratch() {
  int dev = 0;
  DPMI_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto& v = device_ptrs();
  if ((int)v.size() <= dev) v.resize(dev + 1, nullptr);
  if (!v[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipError(std::string("hipGetSymbolAddress of the virial scratch failed: ") + hipGetErrorString(e));
    v[dev] = static_cast<double*>(p);
  }
  return v[dev];
}

}  // namespace dpmi
