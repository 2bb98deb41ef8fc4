// Device context of the C++ API (no reference counterpart): calls on one
// host thread share one xts_ctx, created on first use on the device named by
// the XTS_DEVICE environment variable (default 0).
#pragma once

#include <cstdint>

#include "xtsrcmaps_amd.h"

namespace Fermi {

// The calling thread's context (created on first use).
xts_ctx* xts_context();

// Use a caller-owned context on this thread (nullptr: create one on next use).
void set_xts_context(xts_ctx* ctx);

// Cubature arithmetic of point_src_model_map_wcs and cubature_convolve_map_psf
// on this thread: XTS_COMPUTE_F32 (default; FP64 geometry, separations near
// the source and scaling) or XTS_COMPUTE_F64.
void set_model_map_flags(int32_t flags);

}  // namespace Fermi
