// raster.hip — light-view depth pass and camera visibility/G-buffer pass for gfx950.
//
// Replaces the reference's fixed-function OpenGL rasterization of the light
// view (ShadowMapping/src/main.cpp:221-274 with Scene.vert:11-20, depth into a
// GL_DEPTH_COMPONENT32F texture cleared to 1.0, depth test LESS, no culling)
// and of the camera G-buffer (GBuffer.vert:15-21, GBuffer.frag:35-37).
//
// Design (binned raster, one launch sequence for a whole batch of maps; the
// kernels live in raster_kernels.hpp):
//   1. k_bin      one thread per (triangle, map): transform, classify against
//                 the clip planes, snap, and count the screen tiles each
//                 triangle overlaps (LDS-aggregated); wide and clipped pieces
//                 become setup records (k_clip, k_bin_wide);
//   2. k_scan     one workgroup per map: exclusive scan of the tile counts;
//   3. k_bin      scatter pass: append the triangle id to every overlapped
//                 tile's list;
//   4. k_raster   one workgroup per (tile, map): resolve the tile's list into an
//                 LDS tile with ds_min on the FP32 depth bits (non-negative
//                 floats order like uint32) or on 64-bit (depth, id) keys,
//                 then write the tile to HBM once with 16-byte stores.
// atomicMin makes the result independent of list order, so the output is
// deterministic.  When a map's lists would exceed the workspace, that map falls
// back to every workgroup scanning the whole mesh (slower, same result).
//
// Raster rules (the choices the GL spec leaves to the implementation):
//   clip near (z >= -w), far (z <= w) and a guard band |x|,|y| <= 4w with
//   Sutherland-Hodgman (intersection taken from the inside vertex, so a shared
//   edge clips identically in both triangles), fan triangulation; window
//   coordinates snapped to 1/256 pixel with round-to-nearest-even; pixel
//   centres at (i+.5, j+.5); top-left tie rule in GL's y-up window space;
//   window depth linear in screen space, evaluated as an FP32 plane about
//   vertex 0, z = z0 + (A*(Px-X0) + B*(Py-Y0)), whose slopes are derived in
//   FP64 from the exact integer geometry and rounded once; clamped to [0, 1];
//   optional polygon offset factor*max|dz/dx|,|dz/dy| + units*2^(e-23).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"
#include "gi_internal.hpp"
#include "raster_kernels.hpp"

namespace gi {

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

// Binning of every map (count, clip, scan, scatter) and the raster of the
// maps whose lists overflowed the workspace.
template <bool kVis>
int bin_stage(const MeshView& mesh, const MvpSource& mvps, int nmaps, const RasterCfg& cfg, const Workspace& ws,
              void* out, hipStream_t s) {
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    const BinBufs& b = ws.bins;
    size_t nz = zero_words(nmaps, cfg);
    size_t fb = (nz + 255) / 256;
    unsigned fill_blocks = (unsigned)(fb < 4096 ? fb : 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks), dim3(256), 0, s, b.counts, nz, 0u);
    dim3 gb((mesh.nt + 255) / 256, nmaps);
    dim3 gw(kWideBlocks, nmaps);
    if (mesh.nt > 0) {
        hipLaunchKernelGGL((k_bin<kTile, false>), gb, dim3(256), 0, s, mesh, mvps, cfg, b);
        hipLaunchKernelGGL(k_clip, dim3(kClipBlocks, nmaps), dim3(kClipThreads), 0, s, mesh, mvps, cfg, b);
        hipLaunchKernelGGL((k_bin_wide<kTile, false>), gw, dim3(256), 0, s, cfg, b);
    }
    hipLaunchKernelGGL(k_scan, dim3(nmaps), dim3(kScanThreads), 0, s, b, ntiles);
    if (mesh.nt > 0) {
        if constexpr (wide_span<kTile>() <= 4) {
            // triangles' appends and the wide records' appends in one launch
            const int tri_blocks = (mesh.nt + 256 * kScatterTris - 1) / (256 * kScatterTris);
            hipLaunchKernelGGL((k_bin_scatter<kTile, kScatterTris>), dim3(tri_blocks + kWideBlocks, nmaps), dim3(256),
                               0, s, mesh, cfg, b, tri_blocks);
        } else {
            hipLaunchKernelGGL((k_bin<kTile, true>), gb, dim3(256), 0, s, mesh, mvps, cfg, b);
            hipLaunchKernelGGL((k_bin_wide<kTile, true>), gw, dim3(256), 0, s, cfg, b);
        }
    }
    hipLaunchKernelGGL((k_raster_fallback<kTile, kVis>), dim3(ntiles), dim3(kRasterThreads), 0, s, mesh, mvps, cfg, b,
                       out, nmaps);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

// Tile raster of maps [map_begin, map_begin + map_count) from the lists the
// bin stage left in the workspace (maps that overflowed are skipped: the bin
// stage rastered them).
template <bool kVis>
int raster_stage(const MeshView& mesh, const MvpSource& mvps, int map_begin, int map_count, const RasterCfg& cfg,
                 const Workspace& ws, void* out, hipStream_t s, hipEvent_t ev_begin = nullptr,
                 hipEvent_t ev_end = nullptr) {
    const int ntiles = cfg.tiles_x * cfg.tiles_y;
    if (ev_begin && hipEventRecord(ev_begin, s) != hipSuccess) return GI_EDEVICE;
    hipLaunchKernelGGL((k_raster<kTile, kVis>), dim3((unsigned)(ntiles * map_count)), dim3(kRasterThreads), 0, s, mesh,
                       mvps, cfg, ws.bins, out, map_begin, map_count);
    if (ev_end && hipEventRecord(ev_end, s) != hipSuccess) return GI_EDEVICE;
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

template <bool kVis>
int run_raster(const MeshView& mesh, const MvpSource& mvps, int nmaps, const RasterCfg& cfg, const Workspace& ws,
               void* out, hipStream_t s, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr) {
    int rc = bin_stage<kVis>(mesh, mvps, nmaps, cfg, ws, out, s);
    if (rc != GI_OK) return rc;
    return raster_stage<kVis>(mesh, mvps, 0, nmaps, cfg, ws, out, s, ev_begin, ev_end);
}

}  // namespace

namespace internal {

size_t vis_raster_workspace_size(int ntris, int nmaps, int w, int h) {
    if (ntris < 0 || nmaps <= 0 || !size_ok(w, h)) return 0;
    return layout(ntris, nmaps, w, h, false, nullptr, nullptr);
}

int vis_raster(const float* pos, int nverts, const int32_t* idx, int ntris, const float* mvps_dev, int nmaps, int w,
               int h, unsigned long long* vis_out, void* ws, size_t ws_bytes, hipStream_t s) {
    if (!pos || (!idx && ntris > 0) || !mvps_dev || !vis_out || !ws || nverts <= 0 || ntris < 0 || nmaps <= 0)
        return GI_EINVAL;
    if (ntris >= (1 << 28)) return GI_ENOTSUP;
    if (!size_ok(w, h)) return GI_ENOTSUP;
    Workspace wsp;
    size_t need = layout(ntris, nmaps, w, h, false, &wsp, reinterpret_cast<char*>(ws));
    if (ws_bytes < need) return GI_ENOSPACE;
    MeshView mesh{pos, nullptr, idx, ntris};
    MvpSource mvps{};
    mvps.dev = mvps_dev;
    RasterCfg cfg{w, h, (w + kTile - 1) / kTile, (h + kTile - 1) / kTile, 0.0f, 0.0f};
    return run_raster<true>(mesh, mvps, nmaps, cfg, wsp, vis_out, s);
}

}  // namespace internal
}  // namespace gi

using namespace gi;

extern "C" {

size_t gi_depth_pass_workspace_size(int ntris, int nmaps, int width, int height) {
    if (ntris < 0 || nmaps <= 0 || !size_ok(width, height)) return 0;
    return layout(ntris, nmaps, width, height, false, nullptr, nullptr);
}

}  // extern "C"

namespace {

struct DepthCall {
    MeshView mesh;
    MvpSource mvps;
    RasterCfg cfg;
    Workspace ws;
};

// Argument checks and workspace layout shared by the depth-pass entry points.
int depth_setup(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                float* depth_out_dev, void* workspace_dev, size_t workspace_bytes, DepthCall& d) {
    if (!positions_dev || (!indices_dev && ntris > 0) || !mvps_dev || !depth_out_dev || !workspace_dev)
        return GI_EINVAL;
    if (nverts <= 0 || ntris < 0 || nmaps <= 0) return GI_EINVAL;
    if (ntris >= (1 << 28)) return GI_ENOTSUP;
    if (!size_ok(width, height)) return GI_ENOTSUP;
    if ((reinterpret_cast<uintptr_t>(depth_out_dev) & 15u) != 0) return GI_EINVAL;
    float off_f = opts ? opts->polygon_offset_factor : 0.0f;
    float off_u = opts ? opts->polygon_offset_units : 0.0f;
    if (!std::isfinite(off_f) || !std::isfinite(off_u)) return GI_ENONFINITE;
    size_t need = layout(ntris, nmaps, width, height, false, &d.ws, reinterpret_cast<char*>(workspace_dev));
    if (workspace_bytes < need) return GI_ENOSPACE;
    d.mesh = MeshView{positions_dev, nullptr, indices_dev, ntris};
    d.mvps = MvpSource{};
    d.mvps.dev = mvps_dev;
    d.cfg = RasterCfg{width, height, (width + kTile - 1) / kTile, (height + kTile - 1) / kTile, off_f, off_u};
    return GI_OK;
}

}  // namespace

extern "C" {

int gi_depth_pass(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                  const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                  float* depth_out_dev, void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    DepthCall d;
    int rc = depth_setup(positions_dev, nverts, indices_dev, ntris, mvps_dev, nmaps, width, height, opts,
                         depth_out_dev, workspace_dev, workspace_bytes, d);
    if (rc != GI_OK) return rc;
    return run_raster<false>(d.mesh, d.mvps, nmaps, d.cfg, d.ws, depth_out_dev, reinterpret_cast<hipStream_t>(stream),
                             opts ? reinterpret_cast<hipEvent_t>(opts->raster_begin_event) : nullptr,
                             opts ? reinterpret_cast<hipEvent_t>(opts->raster_end_event) : nullptr);
}

int gi_depth_pass_bin(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                      const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                      float* depth_out_dev, void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    DepthCall d;
    int rc = depth_setup(positions_dev, nverts, indices_dev, ntris, mvps_dev, nmaps, width, height, opts,
                         depth_out_dev, workspace_dev, workspace_bytes, d);
    if (rc != GI_OK) return rc;
    return bin_stage<false>(d.mesh, d.mvps, nmaps, d.cfg, d.ws, depth_out_dev, reinterpret_cast<hipStream_t>(stream));
}

int gi_depth_pass_raster(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                         const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                         int map_begin, int map_count, float* depth_out_dev, void* workspace_dev,
                         size_t workspace_bytes, gi_stream stream) {
    DepthCall d;
    int rc = depth_setup(positions_dev, nverts, indices_dev, ntris, mvps_dev, nmaps, width, height, opts,
                         depth_out_dev, workspace_dev, workspace_bytes, d);
    if (rc != GI_OK) return rc;
    if (map_begin < 0 || map_count <= 0 || map_begin > nmaps - map_count) return GI_EINVAL;
    return raster_stage<false>(d.mesh, d.mvps, map_begin, map_count, d.cfg, d.ws, depth_out_dev,
                               reinterpret_cast<hipStream_t>(stream),
                               opts ? reinterpret_cast<hipEvent_t>(opts->raster_begin_event) : nullptr,
                               opts ? reinterpret_cast<hipEvent_t>(opts->raster_end_event) : nullptr);
}

size_t gi_gbuffer_pass_workspace_size(int ntris, int width, int height) {
    if (ntris < 0 || !size_ok(width, height)) return 0;
    return layout(ntris, 1, width, height, true, nullptr, nullptr);
}

int gi_gbuffer_pass(const float* positions_dev, const float* normals_dev, int nverts, const int32_t* indices_dev,
                    int ntris, const float mvp[16], int width, int height, float* position_out_dev,
                    float* normal_out_dev, float* facing_out_dev, void* workspace_dev, size_t workspace_bytes,
                    gi_stream stream) {
    if (!positions_dev || !normals_dev || (!indices_dev && ntris > 0) || !mvp || !position_out_dev ||
        !normal_out_dev || !workspace_dev)
        return GI_EINVAL;
    if (nverts <= 0 || ntris < 0) return GI_EINVAL;
    if (ntris >= (1 << 28)) return GI_ENOTSUP;
    if (!size_ok(width, height)) return GI_ENOTSUP;
    for (int i = 0; i < 16; ++i)
        if (!std::isfinite(mvp[i])) return GI_ENONFINITE;
    Workspace ws;
    size_t need = layout(ntris, 1, width, height, true, &ws, reinterpret_cast<char*>(workspace_dev));
    if (workspace_bytes < need) return GI_ENOSPACE;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MeshView mesh{positions_dev, normals_dev, indices_dev, ntris};
    MvpSource mvps{};
    mvps.dev = nullptr;
    for (int i = 0; i < 16; ++i) mvps.single.m[i] = mvp[i];
    RasterCfg cfg{width, height, (width + kTile - 1) / kTile, (height + kTile - 1) / kTile, 0.0f, 0.0f};
    int rc = run_raster<true>(mesh, mvps, 1, cfg, ws, ws.vis, s);
    if (rc != GI_OK) return rc;
    size_t npx = (size_t)width * height;
    hipLaunchKernelGGL(k_gbuffer_resolve, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, s, mesh, mvps.single,
                       cfg, reinterpret_cast<const unsigned long long*>(ws.vis), position_out_dev, normal_out_dev,
                       facing_out_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // extern "C"
