// noalloc.hip — the depth pass without a caller workspace, for gfx950.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "../../include/gi.h"
// ---------------------------------------------------------------------------
// Depth pass without a caller workspace (gi_depth_pass_noalloc): for callers
// that cannot allocate device memory (code under graph capture, tight serving
// loops).  The workspace is a file-scope device array, one copy per GPU,
// held until the process ends.  Its device address is looked up once per
// device, with that device current, and only that address is handed to the
// kernels (the array's name is a host-side stand-in).  The binning kernels
// clear the counters they need at the start of every call, so nothing from
// an earlier call is read; calls that overlap on one GPU share the array and
// race.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (nullptr on failure).
void* scratch_on_current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

size_t gi_depth_pass_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int gi_depth_pass_noalloc(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                          const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                          float* depth_out_dev, gi_stream stream) {
    if (ntris < 0 || nmaps <= 0 || width <= 0 || height <= 0 || width > GI_MAX_EDGE || height > GI_MAX_EDGE)
        return GI_EINVAL;
    const size_t need = gi_depth_pass_workspace_size(ntris, nmaps, width, height);
    if (need == 0) return GI_EINVAL;
    if (need > kNoAllocScratchBytes) return GI_ENOSPACE;
    void* ws = scratch_on_current_device();
    if (!ws) return GI_EDEVICE;
    return gi_depth_pass(positions_dev, nverts, indices_dev, ntris, mvps_dev, nmaps, width, height, opts,
                         depth_out_dev, ws, kNoAllocScratchBytes, stream);
}

}  // extern "C"
