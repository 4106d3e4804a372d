// Shared device helpers for the dpmi gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/dpmi.h"

namespace dpmi {

constexpr int kWave = 64;

// Error codes written by kernels into a device-side status word; the host
// turns them into dpmi_status values (include/dpmi.h).
enum DeviceFlag : int {
  kFlagNlistCapacity = 1,     // a neighbour row has more candidates than the LDS row capacity
  kFlagGhostCapacity = 2,     // nall exceeds the extended-coordinate buffer
  kFlagIndexRange = 4,        // an extended index does not fit the 24-bit sort key
  kFlagTypeRange = 8,         // atom type >= 255 (sort key)
  kFlagCellCapacity = 16,     // a cell holds more atoms than the binning assumes
  kFlagHalfRange = 32,        // a mixed-precision matrix-core operand exceeds the FP16 range
  kFlagTypeMismatch = 64,     // a caller's local atom type differs from the engine's
  kFlagGridChanged = 128,     // a box refreshed on the device needs another cell grid
};

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};
// a valid request this build does not implement (DPMI_ERR_UNSUPPORTED)
struct Unsupported : std::runtime_error {
  explicit Unsupported(const std::string& s) : std::runtime_error(s) {}
};

inline void check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    throw HipError(std::string("HIP error in ") + what + " (" + file + ":" + std::to_string(line) +
                   "): " + hipGetErrorString(e));
  }
}
#define DPMI_CHECK(x) ::dpmi::check((x), #x, __FILE__, __LINE__)
#define DPMI_LAUNCH_CHECK() ::dpmi::check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------- device --
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// Global index of the calling wave, marked wave-uniform (readfirstlane) so
// that addresses derived from it go through scalar registers and loads.
__device__ __forceinline__ long long wave_index() {
  return (long long)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// Neumaier-compensated accumulation (sum, compensation).
struct KahanD {
  double s = 0., c = 0.;
  __device__ __forceinline__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

}  // namespace dpmi
