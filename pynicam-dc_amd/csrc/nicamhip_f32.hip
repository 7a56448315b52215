// GPU build of the nicamhip kernels for gfx950, mixed-precision variant:
// fields and stencil arithmetic in FP32, the tridiagonal solve and the global
// sums in FP64 (namespace nh::f32).  Linked with nicamhip.hip.
#define NH_REAL float
#define NH_VARIANT f32

#include "nh_hipexec.h"

using NHX = nh::HipExec;

#include "nh_impl.h"
