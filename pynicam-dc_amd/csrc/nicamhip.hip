// GPU build of the nicamhip C-ABI for gfx950: the FP64 variant of the kernels
// and the extern "C" layer.  The mixed-precision variant is nicamhip_f32.hip;
// both are linked into one library.
#include "nh_hipexec.h"

using NHX = nh::HipExec;

#include "nh_impl.h"
#include "nh_api.h"
