// GPU build of the nicamhip C-ABI: gfx950, FP64 column kernels.
//
// Every functor runs as one thread per column (256-thread workgroups,
// 4 wavefronts); the grid covers all regional and pole columns.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "nh_step.h"
#include "../../include/nicamhip.h"

namespace nh {

template <class F>
__global__ __launch_bounds__(256) void k_run(F f, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f(t);
}

class HipExec {
 public:
  HipExec(int device, void* stream) : s_((hipStream_t)stream) {
    int cur = -1;
    err_ = hipGetDevice(&cur);
    if (err_ == hipSuccess && cur != device) err_ = hipSetDevice(device);
    const char* e = std::getenv("NH_SYNC");
    sync_each_ = (e && e[0] == '1');
  }
  void* alloc(size_t bytes) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      note(e, "hipMalloc");
      return nullptr;
    }
    return p;
  }
  void release(void* p) {
    if (p) (void)hipFree(p);
  }
  void zero(void* p, size_t bytes) { note(hipMemset(p, 0, bytes), "hipMemset"); }
  void h2d(void* dst, const void* src, size_t bytes) {
    note(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void d2d(void* dst, const void* src, size_t bytes) {
    note(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s_), "hipMemcpyAsync D2D");
  }
  template <class F>
  void run(const F& f, long n) {
    if (n <= 0) return;
    const int bs = 256;
    long nb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_run<F>, dim3((unsigned)nb), dim3(bs), 0, s_, f, n);
    note(hipGetLastError(), "kernel launch");
    if (sync_each_) note(hipStreamSynchronize(s_), "kernel execution");
  }
  void sync() { note(hipStreamSynchronize(s_), "hipStreamSynchronize"); }
  int check(const char* what) {
    if (err_ == hipSuccess) note(hipGetLastError(), what);
    return err_ == hipSuccess ? 0 : NH_ERR_DEVICE;
  }
  std::string message() const {
    return std::string(where_) + ": " + hipGetErrorString(err_);
  }

 private:
  void note(hipError_t e, const char* where) {
    if (e != hipSuccess && err_ == hipSuccess) {
      err_ = e;
      where_ = where;
    }
  }
  hipStream_t s_;
  hipError_t err_ = hipSuccess;
  const char* where_ = "";
  bool sync_each_ = false;
};

}  // namespace nh

using NHX = nh::HipExec;

#include "nh_api.h"
