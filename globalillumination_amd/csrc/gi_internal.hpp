// gi_internal.hpp — entry points shared between translation units of libgi
// (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gi::internal {

// Visibility keys ((FP32 depth bits << 32) | tri*8 + fan piece, cleared to ~0)
// of `nmaps` views of one mesh, with the depth pass's raster rules; vis_out
// holds nmaps x h x w keys.  The workspace excludes vis_out.
size_t vis_raster_workspace_size(int ntris, int nmaps, int w, int h);
int vis_raster(const float* pos, int nverts, const int32_t* idx, int ntris, const float* mvps_dev, int nmaps, int w,
               int h, unsigned long long* vis_out, void* ws, size_t ws_bytes, hipStream_t s);

}  // namespace gi::internal
