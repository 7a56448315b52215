// Host build of the mixed-precision variant of the nicamhip kernels (test
// infrastructure), linked with nh_hostemu.cpp.
#define NH_REAL float
#define NH_VARIANT f32

#include "nh_hostexec.h"

using NHX = nh::HostExec;

#include "../../pynicam-dc_amd/csrc/nh_impl.h"
