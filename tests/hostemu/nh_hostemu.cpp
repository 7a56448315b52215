// Host build of the nicamhip library (test infrastructure): the FP64 variant
// of the kernels and the extern "C" layer; nh_hostemu_f32.cpp adds the
// mixed-precision variant.  Same sources as the gfx950 build, NH_HD empty.
#include "nh_hostexec.h"

using NHX = nh::HostExec;

#include "../../pynicam-dc_amd/csrc/nh_impl.h"
#include "../../pynicam-dc_amd/csrc/nh_api.h"

// the allocation-free entry points' scratch on the host: a static array of the
// same size as the gfx950 build's device array
int nh::scratch_lookup(int, double** p, long* n, std::string*) {
  static double host_scratch[1L << 20];
  *p = host_scratch;
  *n = 1L << 20;
  return 0;
}
