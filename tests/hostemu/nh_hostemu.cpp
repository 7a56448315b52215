// Host build of the nicamhip kernels, for CPU tests of the kernel source.
//
// The column functors and the launch sequence are compiled unchanged for the
// host (NH_HD empty); the executor runs every functor as a loop over threads.
// This library is test infrastructure: the product library is the gfx950
// build of csrc/nicamhip.hip.
#define NH_HD
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../pynicam-dc_amd/csrc/nh_step.h"
#include "../../include/nicamhip.h"

namespace nh {

class HostExec {
 public:
  HostExec(int, void*) {}
  void* alloc(size_t bytes) { return std::aligned_alloc(256, (bytes + 255) / 256 * 256); }
  void release(void* p) { std::free(p); }
  void zero(void* p, size_t bytes) { std::memset(p, 0, bytes); }
  void h2d(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); }
  void d2d(void* dst, const void* src, size_t bytes) { std::memmove(dst, src, bytes); }
  template <class F>
  void run(const F& f, long n) {
#pragma omp parallel for schedule(static)
    for (long t = 0; t < n; ++t) f(t);
  }
  void sync() {}
  int check(const char*) { return 0; }
  std::string message() const { return "host"; }
};

}  // namespace nh

using NHX = nh::HostExec;

#include "../../pynicam-dc_amd/csrc/nh_api.h"
