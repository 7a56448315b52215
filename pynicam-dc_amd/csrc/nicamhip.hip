// GPU build of the nicamhip C-ABI for gfx950: the FP64 variant of the kernels
// and the extern "C" layer.  The mixed-precision variant is nicamhip_f32.hip;
// both are linked into one library.
#include "nh_hipexec.h"

using NHX = nh::HipExec;

#include "nh_impl.h"
#include "nh_api.h"

#include <mutex>

// The allocation-free entry points' scratch: a file-scope device array of
// kScratchValues FP64 values (8 MiB), held by every GPU the library runs on
// until the process ends.  In host code the array's name is not a device
// address; its address on a device comes from hipGetSymbolAddress, looked up
// once per device (with that device current) and kept per device.
namespace {
constexpr long kScratchValues = 1L << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kScratchValues];

namespace {
double* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_rc[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];
}  // namespace

int nh::scratch_lookup(int device, double** p, long* n, std::string* err) {
  if (device < 0 || device >= kMaxDevices) {
    if (err) *err = "device index out of range for the scratch table";
    return NH_ERR_ARG;
  }
  std::call_once(g_scratch_once[device], [device] {
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != device) e = hipSetDevice(device);
    void* ptr = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    g_scratch_rc[device] = e;
    g_scratch_ptr[device] = (e == hipSuccess) ? static_cast<double*>(ptr) : nullptr;
  });
  if (g_scratch_rc[device] != hipSuccess) {
    if (err) *err = std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(g_scratch_rc[device]);
    return NH_ERR_DEVICE;
  }
  *p = g_scratch_ptr[device];
  *n = kScratchValues;
  return 0;
}
