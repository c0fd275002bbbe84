// tw_noalloc.hip -- the device scratch array of the allocation-free entry
// point (tw_check_finite_noalloc) and its per-device address.
//
// tw_check_finite allocates a one-word flag with hipMalloc on every call.
// Callers that cannot allocate (code under graph capture, tight serving
// loops, harnesses that forbid it) use the variant that keeps the flag in the
// file-scope array `scratch` below instead.  In host code the name `scratch`
// is only a host-side stand-in for the symbol: its device address is looked
// up once per device (hipGetSymbolAddress, with that device current) and only
// that address is ever passed to kernels or copies.  Each device holds its own
// copy of the array until the process ends; every call on a device shares it,
// so calls that overlap on one device race.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "tw_internal.h"

namespace tw {

// words of scratch per device: the non-finite flag of tw_check_finite_noalloc
// uses word 0; the size does not depend on the grid, so no input exceeds it
constexpr int kScratchWords = 64;  // 256 bytes
static __device__ __align__(256) int32_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int32_t* g_ptr[kMaxDevices] = {};

__global__ void k_fill32(int32_t* p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// Device address of `scratch` on `device` (the current device), resolved on
// the first call for that device and kept.
hipError_t scratch_address(int device, int32_t** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = (int32_t*)p;
  }
  *out = g_ptr[device];
  return hipSuccess;
}

// tw_check_finite's launch with the flag in the device's scratch: clear the
// flag with a fill kernel (no memset node, so the launches stay capturable),
// run the same finiteness kernels on it, and copy the flag to *host_flag.
// The caller synchronizes the stream before reading *host_flag.
hipError_t check_finite_scratch(const Plan& p, int device, hipStream_t s, int32_t* host_flag) {
  int32_t* flag = nullptr;
  hipError_t e = scratch_address(device, &flag);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(64), 0, s, flag, 1, 0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = launch_check_finite(p, flag, s)) != hipSuccess) return e;
  return hipMemcpyAsync(host_flag, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s);
}

}  // namespace tw
