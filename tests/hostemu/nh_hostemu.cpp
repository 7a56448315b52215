// Host build of the nicamhip library (test infrastructure): the FP64 variant
// of the kernels and the extern "C" layer; nh_hostemu_f32.cpp adds the
// mixed-precision variant.  Same sources as the gfx950 build, NH_HD empty.
#include "nh_hostexec.h"

using NHX = nh::HostExec;

#include "../../pynicam-dc_amd/csrc/nh_impl.h"
#include "../../pynicam-dc_amd/csrc/nh_api.h"
