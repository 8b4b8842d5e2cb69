// batch.hip — batched shadow frames for gfx950.
//
// No reference counterpart: the reference renders one light at a time.  A
// batch holds many independent instances (mesh, light samples, map size,
// camera G-buffer, soft-shadow parameters), each the work of one
// renderMonteCarlo frame (SoftShadowMapping/src/main.cpp:756-811).  Instead
// of a launch sequence per instance, every stage of the depth pass runs as one
// persistent kernel over all instances' work items, and the screen pass as one
// persistent kernel per PCF variant:
//   kb_zero, kb_fill       clear the bin counters of every instance and the
//                          work-queue counters;
//   kb_bin, kb_clip,       binning (count), clipping, wide records, scan and
//   kb_bin_wide, kb_scan,  binning (scatter), over (instance, map, triangle
//   kb_bin, kb_bin_wide    block) items in a fixed stride;
//   kb_raster              (instance, map, tile) items from a work queue;
//   kb_raster_fallback     overflowed maps;
//   kb_soft                (instance, pixel block) items from per-XCD work
//                          queues, each XCD draining a contiguous,
//                          cost-balanced range so its L2 serves neighbouring
//                          pixels.
// The per-item work is the single-instance kernels' own device code
// (raster_kernels.hpp, screen_kernels.hpp), so every map and shadow value is
// bit-identical to gi_depth_pass + gi_soft_shadow_pass per instance.  The
// plan's launches can be captured in one hipGraph; gi_batch_autotune picks the
// raster tile edge, the screen block shape and the persistent grid sizes by
// measurement.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/gi.h"
#include "raster_kernels.hpp"
#include "screen_kernels.hpp"

namespace gi {
namespace {

enum Stage { kStFill = 0, kStBin, kStClip, kStWide, kStScan, kStRaster, kStFallback, kStScatter, kStages };
constexpr int kFillChunk = 4096;  // words cleared per fill item
constexpr int kXcds = 8;
constexpr int kScreenThreads = 256;
constexpr int kNtGroups = 3;  // soft variants: generic loop, 3 taps per axis, 4 taps per axis
constexpr int kShapes = 5;    // screen blocks 64x4, 32x8, 16x16 (row-major waves), 16x16 of 8x8 wave
                              // tiles, 64x4 of 16x4 wave tiles
constexpr int kCounters = 64;  // [0] raster queue, [1] a map of the depth run overflowed (until the
                               // screen run clears it), [1 + 8 g + x] screen queue of group g on XCD x,
                               // [63] the autotuner's overflow flag

struct RasterJob {
    MeshView mesh;
    const float* mvps;
    RasterCfg cfg;
    BinBufs b;
    float* out;
    uint32_t nzero;  // words cleared from b.counts on
    int nmaps;
    int tri_blocks;
    int ntiles;
};

struct RasterTable {
    const RasterJob* jobs;
    const uint32_t* prefix;  // kStages x (njobs + 1): items of each stage per job, exclusive prefix
    int njobs;
    uint32_t* counter;       // [0] raster work queue, [1] set when a map overflowed
};

struct SoftJob {
    GBufferView gb;
    const float* maps;
    int smw, smh;
    const float4* trans;
    SoftUniforms u;
    float* out;
    int bx_n;  // pixel blocks per row
};

struct SoftTable {
    const SoftJob* jobs;
    const uint32_t* prefix;  // njobs + 1
    int njobs;
    const uint32_t* xr;      // [2 x, 2 x + 1]: item range of XCD x
    uint32_t* counters;      // one per XCD
};

// Job holding work item `item` (largest j with pre[j] <= item; jobs without
// items share their successor's prefix and are skipped).
__device__ __forceinline__ int find_job(const uint32_t* __restrict__ pre, int njobs, uint32_t item) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ MvpSource mvp_of(const RasterJob& J) {
    MvpSource m{};
    m.dev = J.mvps;
    return m;
}

__global__ void kb_zero(uint32_t* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void kb_fill(RasterTable t) {
    const uint32_t* pre = t.prefix + kStFill * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t w0 = (item - pre[j]) * kFillChunk;
        const uint32_t w1 = min(w0 + (uint32_t)kFillChunk, J.nzero);
        for (uint32_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) J.b.counts[w] = 0u;
    }
}

// A job's binning buffers rebased on the workspace pointer `ws` (a kernel
// argument): the same addresses, but the compiler knows them for device
// memory and addresses them with global instructions.  Pointers read from
// the job table carry no address space, and their flat instructions count
// against the LDS wait counter, so every LDS barrier would also wait for the
// workgroup's outstanding tile-count atomics and record stores.
template <typename T>
__device__ __forceinline__ T* in_ws(char* ws, T* p) {
    return reinterpret_cast<T*>(ws + ((uintptr_t)p - (uintptr_t)ws));
}

__device__ __forceinline__ BinBufs bufs_in_ws(char* ws, const BinBufs& b) {
    BinBufs g = b;
    g.counts = in_ws(ws, b.counts);
    g.offsets = in_ws(ws, b.offsets);
    g.overflow = in_ws(ws, b.overflow);
    g.wide_over = in_ws(ws, b.wide_over);
    g.nwide = in_ws(ws, b.nwide);
    g.wide = in_ws(ws, b.wide);
    g.nclip = in_ws(ws, b.nclip);
    g.clip = in_ws(ws, b.clip);
    g.entries = in_ws(ws, b.entries);
    g.tiles = in_ws(ws, b.tiles);
    g.setups = in_ws(ws, b.setups);
    return g;
}

template <int TS, bool kScatter>
__global__ __launch_bounds__(256) void kb_bin(RasterTable t, char* ws) {
    const uint32_t* pre = t.prefix + kStBin * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t local = item - pre[j];
        bin_body<TS, kScatter>(J.mesh, mvp_of(J), J.cfg, bufs_in_ws(ws, J.b), (int)(local % J.tri_blocks),
                               (int)(local / J.tri_blocks));
        lds_barrier();  // the LDS histogram is reused by the next item
    }
}

// The scatter stage for 64-texel tiles: kScatterTris triangles per thread
// (bin_scatter_body, the single-instance scatter's code), so a workgroup's
// two dependent global round trips cover kScatterTris x 256 triangles.
template <int TS>
__global__ __launch_bounds__(256) void kb_scatter(RasterTable t, char* ws) {
    const uint32_t* pre = t.prefix + kStScatter * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t local = item - pre[j];
        const uint32_t blocks = (uint32_t)(J.tri_blocks + kScatterTris - 1) / kScatterTris;
        bin_scatter_body<TS, kScatterTris>(J.mesh, J.cfg, bufs_in_ws(ws, J.b), (int)(local % blocks),
                                           (int)(local / blocks));
        lds_barrier();  // the LDS histogram is reused by the next item
    }
}

__global__ __launch_bounds__(kClipThreads) void kb_clip(RasterTable t) {
    const uint32_t* pre = t.prefix + kStClip * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t local = item - pre[j];
        clip_body(J.mesh, mvp_of(J), J.cfg, J.b, (int)(local % kClipBlocks), kClipBlocks, (int)(local / kClipBlocks));
    }
}

template <int TS, bool kScatter>
__global__ __launch_bounds__(256) void kb_bin_wide(RasterTable t) {
    const uint32_t* pre = t.prefix + kStWide * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t local = item - pre[j];
        bin_wide_body<TS, kScatter>(J.cfg, J.b, (int)(local % kWideBlocks), kWideBlocks, (int)(local / kWideBlocks));
    }
}

__global__ __launch_bounds__(kScanThreads) void kb_scan(RasterTable t) {
    const uint32_t* pre = t.prefix + kStScan * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const int map = (int)(item - pre[j]);
        scan_body(J.b, J.ntiles, map);
        // the thread that decided the map's overflow raises the batch's flag
        // (kb_raster_fallback does nothing without it)
        if (threadIdx.x == kScanThreads - 1 && J.b.overflow[map]) t.counter[1] = 1u;
        __syncthreads();
    }
}

// flag = 1 when any map of the batch overflowed its tile lists (its raster
// then went through kb_raster_fallback).  One thread per (instance, map).
__global__ void kb_overflow_any(RasterTable t, uint32_t* __restrict__ flag) {
    const uint32_t* pre = t.prefix + kStScan * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x * blockDim.x + threadIdx.x; item < total; item += gridDim.x * blockDim.x) {
        const int j = find_job(pre, t.njobs, item);
        if (t.jobs[j].b.overflow[item - pre[j]]) atomicOr(flag, 1u);
    }
}

// (instance, map, tile) items from a work queue: tiles differ widely in cost.
template <int TS>
__global__ __launch_bounds__(kRasterThreads) void kb_raster(RasterTable t) {
    __shared__ __align__(16) uint32_t tile[TS * TS];
    __shared__ Setup rec[kRecQueue];
    __shared__ int nrec;
    __shared__ uint32_t s_item;
    const uint32_t* pre = t.prefix + kStRaster * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(t.counter, 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= total) break;  // uniform: every workgroup ends once the queue is drained
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const uint32_t local = item - pre[j];
        const int map = (int)(local / J.ntiles), tile_id = (int)(local % J.ntiles);
        if (J.b.overflow[map]) continue;  // left to kb_raster_fallback
        raster_tile<TS, false, false>(J.mesh, mvp_of(J), J.cfg, J.b, J.out, map, tile_id, tile, rec, nrec);
    }
}

template <int TS>
__global__ __launch_bounds__(kRasterThreads) void kb_raster_fallback(RasterTable t) {
    __shared__ uint32_t tile[TS * TS];
    __shared__ Setup rec[kRecQueue];
    __shared__ int nrec;
    if (t.counter[1] == 0u) return;  // no map overflowed (kb_scan)
    const uint32_t* pre = t.prefix + kStFallback * (t.njobs + 1);
    const uint32_t total = pre[t.njobs];
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const int j = find_job(pre, t.njobs, item);
        const RasterJob& J = t.jobs[j];
        const int tile_id = (int)(item - pre[j]);
        for (int map = 0; map < J.nmaps; ++map) {
            if (!J.b.overflow[map]) continue;
            raster_tile<TS, false, true>(J.mesh, mvp_of(J), J.cfg, J.b, J.out, map, tile_id, tile, rec, nrec);
            __syncthreads();
        }
    }
}

// (instance, pixel block) items.  XCD x (hardware hands consecutive
// workgroups to the XCDs round-robin) drains items [xr[2x], xr[2x+1]) through
// its own queue.
template <int NT, int BX, int BY, bool kWaveTiles>
__global__ __launch_bounds__(kScreenThreads) void kb_soft(SoftTable t) {
    static_assert(BX * BY == kScreenThreads, "block shape");
    __shared__ uint32_t s_item;
    const int xcd = (int)(blockIdx.x % kXcds);
    const uint32_t lo = t.xr[2 * xcd], hi = t.xr[2 * xcd + 1];
    int tx = (int)threadIdx.x % BX, ty = (int)threadIdx.x / BX;
    if constexpr (kWaveTiles && BX == 64) {  // four 16x4 wave tiles side by side
        tx = ((int)threadIdx.x >> 6) * 16 + ((int)threadIdx.x & 15);
        ty = ((int)threadIdx.x & 63) >> 4;
    } else if constexpr (kWaveTiles) {
        wave_tile_xy((int)threadIdx.x, tx, ty);
    }
    for (;;) {
        if (threadIdx.x == 0) s_item = lo + atomicAdd(&t.counters[xcd], 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= hi) break;  // uniform: every workgroup ends once its range is drained
        const int j = find_job(t.prefix, t.njobs, item);
        const SoftJob& J = t.jobs[j];
        const uint32_t local = item - t.prefix[j];
        const int x = (int)(local % (uint32_t)J.bx_n) * BX + tx, y = (int)(local / (uint32_t)J.bx_n) * BY + ty;
        if (x < J.gb.w && y < J.gb.h) soft_pixel<NT>(J.gb, J.maps, J.smw, J.smh, J.trans, J.u, J.out, x, y);
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

const int kShapeX[kShapes] = {64, 32, 16, 16, 64};
const int kShapeY[kShapes] = {4, 8, 16, 16, 4};
const int kGroupNt[kNtGroups] = {0, 3, 4};
const int kTiles[2] = {32, 64};

int nt_group(int nt) { return nt == 3 ? 1 : (nt == 4 ? 2 : 0); }

struct Layout {
    size_t counters, rjobs[2], rprefix[2], sjobs[kShapes][kNtGroups], sprefix[kShapes][kNtGroups],
        sxr[kShapes][kNtGroups];
    std::vector<size_t> bins;
    size_t total;
};

bool inst_ok(const gi_batch_instance& I) {
    if (!I.positions_dev || I.nverts <= 0 || I.ntris < 0 || I.ntris >= (1 << 28) || (!I.indices_dev && I.ntris > 0))
        return false;
    if (!I.light_mvps_dev || I.nsamples <= 0 || I.nsamples > GI_MAX_SAMPLES) return false;
    if (!size_ok(I.sm_width, I.sm_height) || !I.maps_dev) return false;
    if ((reinterpret_cast<uintptr_t>(I.maps_dev) & 15u) != 0) return false;
    return true;
}

Layout plan_layout(const gi_batch_instance* inst, int n) {
    Layout L;
    size_t off = 0;
    L.counters = off;
    off = align256(off + sizeof(uint32_t) * kCounters);
    for (int v = 0; v < 2; ++v) {
        L.rjobs[v] = off;
        off = align256(off + sizeof(RasterJob) * (size_t)n);
        L.rprefix[v] = off;
        off = align256(off + sizeof(uint32_t) * kStages * (size_t)(n + 1));
    }
    for (int sh = 0; sh < kShapes; ++sh)
        for (int g = 0; g < kNtGroups; ++g) {
            L.sjobs[sh][g] = off;
            off = align256(off + sizeof(SoftJob) * (size_t)n);
            L.sprefix[sh][g] = off;
            off = align256(off + sizeof(uint32_t) * (size_t)(n + 1));
            L.sxr[sh][g] = off;
            off = align256(off + sizeof(uint32_t) * 2 * kXcds);
        }
    L.bins.resize(n);
    for (int i = 0; i < n; ++i) {
        L.bins[i] = off;
        // one region serves both tile edges: the 32-texel grid has more tiles,
        // the 64-texel one keeps a setup record per (triangle, map)
        size_t need = 0;
        for (int v = 0; v < 2; ++v) {
            const size_t sz = layout(inst[i].ntris, inst[i].nsamples, inst[i].sm_width, inst[i].sm_height, false,
                                     nullptr, nullptr, kTiles[v]);
            if (sz > need) need = sz;
        }
        off = align256(off + need);
    }
    L.total = off;
    return L;
}

struct RasterVariant {
    const RasterJob* jobs;
    const uint32_t* prefix;
    uint32_t total[kStages];
};

struct SoftGroup {
    const SoftJob* jobs;
    const uint32_t* prefix;
    const uint32_t* xr;
    int njobs;
};

}  // namespace
}  // namespace gi

using namespace gi;

struct gi_batch {
    int n;
    int ncu;
    gi_batch_tuning tuning;
    RasterVariant rv[2];
    SoftGroup sg[kShapes][kNtGroups];
    uint32_t* counters;
    char* ws;  // the workspace: every job's binning buffers lie in it
};

struct gi_batch_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace gi {
namespace {

bool tuning_ok(const gi_batch_tuning& t) {
    return (t.tile == 32 || t.tile == 64) && t.raster_groups >= 1 && t.raster_groups <= 8 && t.screen_shape >= 0 &&
           t.screen_shape < kShapes && t.screen_groups >= 1 && t.screen_groups <= 8;
}

unsigned grid_of(uint32_t total, unsigned cap) { return total == 0 ? 0u : (total < cap ? total : cap); }

template <int TS>
void launch_scatter(const RasterVariant& R, const RasterTable& tab, char* ws, unsigned cap, hipStream_t s) {
    unsigned g;
    if constexpr (wide_span<TS>() <= 4) {  // the count pass stored the tile codes
        if ((g = grid_of(R.total[kStScatter], cap)))
            hipLaunchKernelGGL((kb_scatter<TS>), dim3(g), dim3(256), 0, s, tab, ws);
    } else {
        if ((g = grid_of(R.total[kStBin], cap)))
            hipLaunchKernelGGL((kb_bin<TS, true>), dim3(g), dim3(256), 0, s, tab, ws);
    }
}

int run_depth(gi_batch* b, const gi_batch_tuning& T, hipStream_t s) {
    const int v = T.tile == 32 ? 0 : 1;
    const RasterVariant& R = b->rv[v];
    RasterTable tab{R.jobs, R.prefix, b->n, b->counters};
    const unsigned cap = (unsigned)(b->ncu * T.raster_groups);
    hipLaunchKernelGGL(kb_zero, dim3(1), dim3(kCounters), 0, s, b->counters, 2);
    unsigned g;
    if ((g = grid_of(R.total[kStFill], cap))) hipLaunchKernelGGL(kb_fill, dim3(g), dim3(256), 0, s, tab);
#define GI_DEPTH_STAGES(TS)                                                                                        \
    if ((g = grid_of(R.total[kStBin], cap))) hipLaunchKernelGGL((kb_bin<TS, false>), dim3(g), dim3(256), 0, s, tab, b->ws); \
    if ((g = grid_of(R.total[kStClip], cap)))                                                                       \
        hipLaunchKernelGGL(kb_clip, dim3(g), dim3(kClipThreads), 0, s, tab);                                        \
    if ((g = grid_of(R.total[kStWide], cap)))                                                                       \
        hipLaunchKernelGGL((kb_bin_wide<TS, false>), dim3(g), dim3(256), 0, s, tab);                                \
    if ((g = grid_of(R.total[kStScan], cap))) hipLaunchKernelGGL(kb_scan, dim3(g), dim3(kScanThreads), 0, s, tab);   \
    launch_scatter<TS>(R, tab, b->ws, cap, s);                                                                     \
    if ((g = grid_of(R.total[kStWide], cap)))                                                                       \
        hipLaunchKernelGGL((kb_bin_wide<TS, true>), dim3(g), dim3(256), 0, s, tab);                                 \
    if ((g = grid_of(R.total[kStRaster], cap)))                                                                     \
        hipLaunchKernelGGL((kb_raster<TS>), dim3(g), dim3(kRasterThreads), 0, s, tab);                              \
    if ((g = grid_of(R.total[kStFallback], cap)))                                                                   \
        hipLaunchKernelGGL((kb_raster_fallback<TS>), dim3(g), dim3(kRasterThreads), 0, s, tab);
    if (T.tile == 32) {
        GI_DEPTH_STAGES(32)
    } else {
        GI_DEPTH_STAGES(64)
    }
#undef GI_DEPTH_STAGES
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

template <int NT>
void launch_soft(int shape, unsigned grid, const SoftTable& tab, hipStream_t s) {
    switch (shape) {
    case 0: hipLaunchKernelGGL((kb_soft<NT, 64, 4, false>), dim3(grid), dim3(kScreenThreads), 0, s, tab); break;
    case 1: hipLaunchKernelGGL((kb_soft<NT, 32, 8, false>), dim3(grid), dim3(kScreenThreads), 0, s, tab); break;
    case 2: hipLaunchKernelGGL((kb_soft<NT, 16, 16, false>), dim3(grid), dim3(kScreenThreads), 0, s, tab); break;
    case 3: hipLaunchKernelGGL((kb_soft<NT, 16, 16, true>), dim3(grid), dim3(kScreenThreads), 0, s, tab); break;
    default: hipLaunchKernelGGL((kb_soft<NT, 64, 4, true>), dim3(grid), dim3(kScreenThreads), 0, s, tab); break;
    }
}

int run_screen(gi_batch* b, const gi_batch_tuning& T, hipStream_t s) {
    hipLaunchKernelGGL(kb_zero, dim3(1), dim3(kCounters), 0, s, b->counters + 1, kCounters - 1);
    unsigned grid = (unsigned)(b->ncu * T.screen_groups);
    grid = (grid + kXcds - 1) / kXcds * kXcds;
    for (int g = 0; g < kNtGroups; ++g) {
        const SoftGroup& G = b->sg[T.screen_shape][g];
        if (G.njobs == 0) continue;
        SoftTable tab{G.jobs, G.prefix, G.njobs, G.xr, b->counters + 1 + kXcds * g};
        switch (kGroupNt[g]) {
        case 3: launch_soft<3>(T.screen_shape, grid, tab, s); break;
        case 4: launch_soft<4>(T.screen_shape, grid, tab, s); break;
        default: launch_soft<0>(T.screen_shape, grid, tab, s); break;
        }
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

// Whether the last depth run of tuning T overflowed any map's tile lists.
int overflowed(gi_batch* b, const gi_batch_tuning& T, hipStream_t s, bool* any) {
    const RasterVariant& R = b->rv[T.tile == 32 ? 0 : 1];
    RasterTable tab{R.jobs, R.prefix, b->n, b->counters};
    uint32_t* flag = b->counters + kCounters - 1;  // not used by the screen queues of the groups
    hipLaunchKernelGGL(kb_zero, dim3(1), dim3(kCounters), 0, s, flag, 1);
    hipLaunchKernelGGL(kb_overflow_any, dim3(64), dim3(256), 0, s, tab, flag);
    uint32_t h = 0;
    if (hipMemcpyAsync(&h, flag, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) return GI_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return GI_EDEVICE;
    *any = h != 0;
    return GI_OK;
}

int timed(gi_batch* b, const gi_batch_tuning& T, hipStream_t s, int reps, bool depth, float* ms) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return GI_EDEVICE;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return GI_EDEVICE;
    }
    int rc = depth ? run_depth(b, T, s) : run_screen(b, T, s);  // warm-up
    if (depth && rc == GI_OK) {
        // a tiling whose tile lists overflow takes the slow fallback raster:
        // it is not a candidate
        bool any = false;
        rc = overflowed(b, T, s, &any);
        if (rc == GI_OK && any) {
            hipEventDestroy(e0);
            hipEventDestroy(e1);
            *ms = 1e30f;
            return GI_OK;
        }
    }
    if (rc == GI_OK && hipEventRecord(e0, s) != hipSuccess) rc = GI_EDEVICE;
    for (int r = 0; r < reps && rc == GI_OK; ++r) rc = depth ? run_depth(b, T, s) : run_screen(b, T, s);
    if (rc == GI_OK && hipEventRecord(e1, s) != hipSuccess) rc = GI_EDEVICE;
    if (rc == GI_OK && hipEventSynchronize(e1) != hipSuccess) rc = GI_EDEVICE;
    float t = 0.0f;
    if (rc == GI_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = GI_EDEVICE;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t / (float)reps;
    return rc;
}

}  // namespace
}  // namespace gi

extern "C" {

size_t gi_batch_workspace_size(const gi_batch_instance* inst, int n) {
    if (!inst || n <= 0 || n > GI_MAX_BATCH) return 0;
    for (int i = 0; i < n; ++i)
        if (!inst_ok(inst[i])) return 0;
    return plan_layout(inst, n).total;
}

int gi_batch_create(const gi_batch_instance* inst, int n, void* workspace_dev, size_t workspace_bytes,
                    gi_batch** out) {
    if (!out) return GI_EINVAL;
    *out = nullptr;
    if (!inst || n <= 0 || n > GI_MAX_BATCH || !workspace_dev) return GI_EINVAL;
    std::vector<SoftUniforms> su(n);
    for (int i = 0; i < n; ++i) {
        const gi_batch_instance& I = inst[i];
        if (!inst_ok(I)) return GI_EINVAL;
        if (!std::isfinite(I.polygon_offset_factor) || !std::isfinite(I.polygon_offset_units)) return GI_ENONFINITE;
        int rc = soft_setup(&I.gbuffer, I.maps_dev, I.nsamples, I.sm_width, I.sm_height, I.light_mvp, I.trans_dev,
                            &I.soft, I.out_dev, su[i]);
        if (rc != GI_OK) return rc;
    }
    const Layout L = plan_layout(inst, n);
    if (workspace_bytes < L.total) return GI_ENOSPACE;
    char* base = reinterpret_cast<char*>(workspace_dev);
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        return GI_EDEVICE;

    // Host images of the tables, uploaded once.
    std::vector<char> img(L.bins.empty() ? L.total : L.bins[0], 0);
    gi_batch* b = new (std::nothrow) gi_batch;
    if (!b) return GI_ENOSPACE;
    b->n = n;
    b->ncu = ncu;
    b->tuning = gi_batch_tuning{64, 4, 3, 4};
    b->counters = reinterpret_cast<uint32_t*>(base + L.counters);
    b->ws = base;
    for (int v = 0; v < 2; ++v) {
        const int ts = kTiles[v];
        RasterJob* jobs = reinterpret_cast<RasterJob*>(img.data() + L.rjobs[v]);
        uint32_t* pre = reinterpret_cast<uint32_t*>(img.data() + L.rprefix[v]);
        uint32_t acc[kStages] = {};
        for (int i = 0; i < n; ++i) {
            const gi_batch_instance& I = inst[i];
            Workspace ws;
            layout(I.ntris, I.nsamples, I.sm_width, I.sm_height, false, &ws, base + L.bins[i], ts);
            RasterJob& J = jobs[i];
            J.mesh = MeshView{I.positions_dev, nullptr, I.indices_dev, I.ntris};
            J.mvps = I.light_mvps_dev;
            J.cfg = RasterCfg{I.sm_width, I.sm_height, (I.sm_width + ts - 1) / ts, (I.sm_height + ts - 1) / ts,
                              I.polygon_offset_factor, I.polygon_offset_units};
            J.b = ws.bins;
            J.out = I.maps_dev;
            J.nzero = (uint32_t)zero_words(I.nsamples, J.cfg);
            J.nmaps = I.nsamples;
            J.tri_blocks = (I.ntris + 255) / 256;
            J.ntiles = J.cfg.tiles_x * J.cfg.tiles_y;
            const uint32_t items[kStages] = {
                cdiv(J.nzero, kFillChunk),
                I.ntris > 0 ? (uint32_t)(J.nmaps * J.tri_blocks) : 0u,
                I.ntris > 0 ? (uint32_t)(J.nmaps * kClipBlocks) : 0u,
                I.ntris > 0 ? (uint32_t)(J.nmaps * kWideBlocks) : 0u,
                (uint32_t)J.nmaps,
                (uint32_t)(J.nmaps * J.ntiles),
                (uint32_t)J.ntiles,
                I.ntris > 0 ? (uint32_t)(J.nmaps * cdiv((uint32_t)J.tri_blocks, (uint32_t)kScatterTris)) : 0u};
            for (int st = 0; st < kStages; ++st) {
                pre[st * (n + 1) + i] = acc[st];
                acc[st] += items[st];
            }
        }
        for (int st = 0; st < kStages; ++st) {
            pre[st * (n + 1) + n] = acc[st];
            b->rv[v].total[st] = acc[st];
        }
        b->rv[v].jobs = reinterpret_cast<const RasterJob*>(base + L.rjobs[v]);
        b->rv[v].prefix = reinterpret_cast<const uint32_t*>(base + L.rprefix[v]);
    }
    for (int sh = 0; sh < kShapes; ++sh) {
        for (int g = 0; g < kNtGroups; ++g) {
            SoftJob* jobs = reinterpret_cast<SoftJob*>(img.data() + L.sjobs[sh][g]);
            uint32_t* pre = reinterpret_cast<uint32_t*>(img.data() + L.sprefix[sh][g]);
            uint32_t* xr = reinterpret_cast<uint32_t*>(img.data() + L.sxr[sh][g]);
            std::vector<double> cost;  // per job: cost of one block
            int nj = 0;
            uint32_t acc = 0;
            for (int i = 0; i < n; ++i) {
                if (nt_group(soft_variant(su[i])) != g) continue;
                const gi_batch_instance& I = inst[i];
                SoftJob& J = jobs[nj];
                J.gb = GBufferView{I.gbuffer.position_dev, nullptr, I.gbuffer.width, I.gbuffer.height};
                J.maps = I.maps_dev;
                J.smw = I.sm_width;
                J.smh = I.sm_height;
                J.trans = reinterpret_cast<const float4*>(I.trans_dev);
                J.u = su[i];
                J.out = I.out_dev;
                J.bx_n = (int)cdiv((uint32_t)I.gbuffer.width, (uint32_t)kShapeX[sh]);
                pre[nj] = acc;
                acc += (uint32_t)J.bx_n * cdiv((uint32_t)I.gbuffer.height, (uint32_t)kShapeY[sh]);
                const double taps = su[i].pcf ? (double)su[i].ntaps * su[i].ntaps : 1.0;
                cost.push_back((double)I.nsamples * taps);
                ++nj;
            }
            pre[nj] = acc;
            // XCD ranges: contiguous item ranges of equal total cost
            double total = 0.0;
            for (int j = 0; j < nj; ++j) total += cost[j] * (double)(pre[j + 1] - pre[j]);
            for (int x = 0; x <= kXcds; ++x) {
                const double target = total * x / kXcds;
                uint32_t item = acc;
                double run = 0.0;
                for (int j = 0; j < nj; ++j) {
                    const double c = cost[j] * (double)(pre[j + 1] - pre[j]);
                    if (run + c >= target) {
                        double k = std::ceil((target - run) / cost[j]);
                        uint32_t off = (uint32_t)(k < 0.0 ? 0.0 : k);
                        item = pre[j] + (off > pre[j + 1] - pre[j] ? pre[j + 1] - pre[j] : off);
                        break;
                    }
                    run += c;
                }
                if (x == 0) item = 0;
                if (x == kXcds) item = acc;
                if (x > 0) xr[2 * (x - 1) + 1] = item;
                if (x < kXcds) xr[2 * x] = item;
            }
            b->sg[sh][g] = SoftGroup{reinterpret_cast<const SoftJob*>(base + L.sjobs[sh][g]),
                                     reinterpret_cast<const uint32_t*>(base + L.sprefix[sh][g]),
                                     reinterpret_cast<const uint32_t*>(base + L.sxr[sh][g]), nj};
        }
    }
    if (hipMemcpy(base, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess) {
        delete b;
        return GI_EDEVICE;
    }
    *out = b;
    return GI_OK;
}

int gi_batch_set_tuning(gi_batch* b, const gi_batch_tuning* t) {
    if (!b || !t || !tuning_ok(*t)) return GI_EINVAL;
    b->tuning = *t;
    return GI_OK;
}

int gi_batch_get_tuning(const gi_batch* b, gi_batch_tuning* t) {
    if (!b || !t) return GI_EINVAL;
    *t = b->tuning;
    return GI_OK;
}

int gi_batch_run(gi_batch* b, gi_stream stream) {
    if (!b) return GI_EINVAL;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int rc = run_depth(b, b->tuning, s);
    return rc == GI_OK ? run_screen(b, b->tuning, s) : rc;
}

int gi_batch_autotune(gi_batch* b, gi_stream stream, int reps, gi_batch_tuning* best, float* best_ms) {
    if (!b || reps <= 0 || reps > 1000) return GI_EINVAL;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int groups[4] = {1, 2, 4, 8};
    gi_batch_tuning T = b->tuning, bestT = b->tuning;
    float best_r = 1e30f, best_s = 1e30f;
    for (int v = 0; v < 2; ++v)
        for (int k = 0; k < 4; ++k) {
            T.tile = kTiles[v];
            T.raster_groups = groups[k];
            float ms = 0.0f;
            int rc = timed(b, T, s, reps, true, &ms);
            if (rc != GI_OK) return rc;
            if (ms < best_r) {
                best_r = ms;
                bestT.tile = T.tile;
                bestT.raster_groups = T.raster_groups;
            }
        }
    // the maps of the last depth run stay in place for the screen candidates
    for (int sh = 0; sh < kShapes; ++sh)
        for (int k = 0; k < 4; ++k) {
            T.screen_shape = sh;
            T.screen_groups = groups[k];
            float ms = 0.0f;
            int rc = timed(b, T, s, reps, false, &ms);
            if (rc != GI_OK) return rc;
            if (ms < best_s) {
                best_s = ms;
                bestT.screen_shape = sh;
                bestT.screen_groups = groups[k];
            }
        }
    b->tuning = bestT;
    if (best) *best = bestT;
    if (best_ms) *best_ms = best_r + best_s;
    return GI_OK;
}

int gi_batch_graph_create(gi_batch* b, gi_batch_graph** out) {
    if (!b || !out) return GI_EINVAL;
    *out = nullptr;
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return GI_EDEVICE;
    gi_batch_graph* g = new (std::nothrow) gi_batch_graph{nullptr, nullptr};
    if (!g) {
        hipStreamDestroy(cs);
        return GI_ENOSPACE;
    }
    int rc = GI_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = GI_EDEVICE;
    if (rc == GI_OK) {
        int rrc = gi_batch_run(b, cs);
        if (hipStreamEndCapture(cs, &g->graph) != hipSuccess || rrc != GI_OK) rc = rrc != GI_OK ? rrc : GI_EDEVICE;
    }
    if (rc == GI_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) rc = GI_EDEVICE;
    hipStreamDestroy(cs);
    if (rc != GI_OK) {
        gi_batch_graph_destroy(g);
        return rc;
    }
    *out = g;
    return GI_OK;
}

int gi_batch_graph_launch(gi_batch_graph* g, gi_stream stream) {
    if (!g || !g->exec) return GI_EINVAL;
    return hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)) == hipSuccess ? GI_OK : GI_EDEVICE;
}

void gi_batch_graph_destroy(gi_batch_graph* g) {
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
}

void gi_batch_destroy(gi_batch* b) { delete b; }

}  // extern "C"
