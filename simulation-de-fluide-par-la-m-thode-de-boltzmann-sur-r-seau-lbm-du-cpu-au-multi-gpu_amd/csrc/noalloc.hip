// lbm_macroscopic_noalloc: lbm_macroscopic for callers that cannot allocate device memory
// (code that must not call hipMalloc once set up: tight field-output loops, harnesses that
// forbid allocation). lbm_macroscopic stages the three fields in a device buffer that the
// handle allocates on its first call (and re-allocates for a larger request); this variant
// stages them in `scratch`, a fixed device array of this library, instead, and calls no
// hipMalloc, hipFree or other allocating routine.
//
// The array is a file-scope __device__ variable: every GPU that the library runs on holds its
// own copy, at its own address, until the process ends. In host code the name `scratch` is a
// host-side stand-in, not a device address, so it is never passed to a kernel or a copy: the
// device address is looked up once per device with hipGetSymbolAddress (with that device
// current) and kept in scratch_ptr[device].
//
// Limits and sharing: 3 * nx * ny reals must fit in kScratchBytes (12 MiB: 1,048,576 cells in
// float and FP16 storage, 524,288 in double); larger grids are rejected with LBM_EINVAL. All
// calls on one GPU share its array, which keeps the last call's fields; nothing in it needs to
// start at zero (the fields kernel writes every cell). Calls from this process are serialised
// by a lock held until the copies to the caller's arrays have completed; calls that overlapped
// on one GPU without it (another copy of the library, or a caller that runs the kernels on
// `scratch` itself) would race.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <string>

#include "host_common.hpp"
#include "lbm_hip.h"
#include "solver_impl.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(12) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;                          // guards scratch_ptr, serialises the calls
std::array<void*, kMaxDevices> scratch_ptr = {};   // device address of `scratch`, per device

// The device address of `scratch` on `device`, looked up on the first call for that device
// (with it current; the caller's current device is restored). Called with scratch_mutex held.
int scratch_on(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return lbmhip::fail(LBM_EINVAL, "device index beyond the no-allocation scratch table");
  if (!scratch_ptr[device]) {
    int current = 0;
    LBM_HIP_TRY(hipGetDevice(&current));
    if (current != device) LBM_HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    const hipError_t back = current != device ? hipSetDevice(current) : hipSuccess;
    if (e != hipSuccess || !ptr)
      return lbmhip::fail(LBM_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
    if (back != hipSuccess)
      return lbmhip::fail(LBM_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(back));
    scratch_ptr[device] = ptr;
  }
  *out = scratch_ptr[device];
  return LBM_OK;
}

}  // namespace

extern "C" int lbm_macroscopic_noalloc(lbm_solver* s, void* rho, void* ux, void* uy) {
  if (!s) return lbmhip::fail(LBM_EINVAL, "null solver handle");
  if (!rho || !ux || !uy) return lbmhip::fail(LBM_EINVAL, "null output array");
  const size_t need = 3 * s->cells() * lbmhip::real_size(s->prec);
  if (need > kScratchBytes)
    return lbmhip::fail(LBM_EINVAL,
                        "lbm_macroscopic_noalloc: the grid needs " + std::to_string(need) +
                            " bytes of scratch (3 fields of nx * ny reals), the maximum is " +
                            std::to_string(kScratchBytes) +
                            " (1,048,576 cells in float or FP16 storage, 524,288 in double)");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  void* ptr = nullptr;
  LBM_TRY(scratch_on(s->device, &ptr));
  return s->macroscopic_into(ptr, rho, ux, uy);
}
