// screen.hip — full-screen shadow passes for gfx950.
//
// One thread per view pixel; the G-buffer is read as SoA planes so each wave
// loads 64 consecutive floats per plane (coalesced), and the shadow map is
// gathered through the vector L1/L2 path.  Per-pixel arithmetic follows the
// reference fragment shaders op for op (FP32, no contraction):
//   Shadow.frag:222-273 (pre-evaluation, naive, PCF),
//   RBSM/NonConservativeSMSR.frag:23-401 (binary RBSM, RPCF),
//   RBSM/FilteredRBSM.frag:234-402 (fractional RBSM),
//   RBSM/ConservativeSMSR.frag:23-235 (conservative RBSM, RPCF),
//   SoftShadow/AccurateSoftShadow.frag:52-133 (area light),
//   SoftShadow/PlausibleSoftShadow.frag:166-632 (PCSS),
//   GBuffer/PhongShading.frag:11-46 (deferred shading),
//   SoftShadow/RBSSM.frag (revectorization-based soft shadows),
//   HardShadow/Discontinuity.frag, QuadTree/RevectorizationBasedReprojection.frag
//   and SoftShadow/RevectorizationBasedAccurateSoftShadow.frag (revectorization-
//   based adaptive sampling).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>

#include "../../include/gi.h"
#include "gi_device.hpp"
#include "screen_kernels.hpp"
#include "rbssm_kernels.hpp"

namespace gi {
namespace {

bool finite_all(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

bool gbuffer_ok(const gi_gbuffer* g) {
    return g && g->position_dev && g->width > 0 && g->height > 0 && g->width <= GI_MAX_EDGE &&
           g->height <= GI_MAX_EDGE;
}

dim3 grid_for(int w, int h) { return dim3((w + kBlockX - 1) / kBlockX, (h + kBlockY - 1) / kBlockY); }

// CUs of the current device (looked up once per device).
int device_cus() {
    static std::atomic<int> cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Workgroups a list-expanding kernel gets: twice as many as the GPU holds at
// once.  Its waves take fixed shares of the list, and the second half of the
// workgroups starts on the CUs whose first workgroups drew cheap entries, a
// balance at workgroup grain that costs no atomics.  On a 1920x1080 view of a
// 200k-triangle soup, RPCF's taps (3 per axis) take 0.167 ms at twice the
// resident workgroups against 0.191 ms at once and 0.164 ms at four times (9
// taps: 0.125, 0.137 and 0.139 ms).  Looked up once per kernel (one gfx950
// code object: the same on every device).
template <auto kKernel>
unsigned expand_blocks() {
    static std::atomic<int> cached{0};
    int per_cu = cached.load(std::memory_order_relaxed);
    if (per_cu <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kKernel, kBlockThreads, 0) != hipSuccess ||
            per_cu <= 0)
            per_cu = 4;
        cached.store(per_cu, std::memory_order_relaxed);
    }
    return (unsigned)(device_cus() * per_cu * 2);
}

}  // namespace
}  // namespace gi

using namespace gi;

extern "C" {

size_t gi_hard_shadow_workspace_size(int width, int height) {
    if (width <= 0 || height <= 0 || width > GI_MAX_EDGE || height > GI_MAX_EDGE) return 0;
    return hard_list_bytes(width, height);
}

int gi_hard_shadow_pass(const gi_gbuffer* gbuffer, const float* depth_dev, const gi_shadow_params* p,
                        float* out_dev, void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !depth_dev || !p || !out_dev) return GI_EINVAL;
    if (p->technique < GI_NAIVE || p->technique > GI_RPCF_CONSERVATIVE) return GI_EINVAL;
    if (p->shadow_map_width <= 0 || p->shadow_map_height <= 0 || p->shadow_map_width > GI_MAX_EDGE ||
        p->shadow_map_height > GI_MAX_EDGE)
        return GI_EINVAL;
    if (!finite_all(p->light_mvp, 16) || !std::isfinite(p->shadow_intensity) || !std::isfinite(p->depth_threshold))
        return GI_ENONFINITE;
    bool uses_loop = p->technique == GI_PCF || p->technique == GI_RPCF || p->technique == GI_RBSM_FILTERED ||
                     p->technique == GI_RPCF_CONSERVATIVE;
    // The PCF loops step by 2*penumbra/kernelOrder: a zero penumbra would never
    // advance, so it is rejected along with orders past the tap budget.
    if (uses_loop && (p->kernel_order <= 0 || p->penumbra_size <= 0 || p->kernel_order > 64 ||
                      p->penumbra_size > 64))
        return GI_EINVAL;
    if (p->technique >= GI_RBSM && (p->max_search < 0 || p->max_search > 4096)) return GI_EINVAL;
    const bool listed =
        p->technique == GI_RPCF || p->technique == GI_RBSM_FILTERED || p->technique == GI_RPCF_CONSERVATIVE;
    if (listed) {
        if (!workspace_dev || (reinterpret_cast<uintptr_t>(workspace_dev) & 15u) != 0) return GI_EINVAL;
        if (workspace_bytes < hard_list_bytes(gbuffer->width, gbuffer->height)) return GI_ENOSPACE;
    }
    if (p->pre_evaluation) {
        if (!gbuffer->normal_dev) return GI_EINVAL;
        if (!finite_all(p->mv, 16) || !finite_all(p->normal_matrix, 9) || !finite_all(p->light_position, 3))
            return GI_ENONFINITE;
    }
    HardUniforms u;
    for (int i = 0; i < 16; ++i) { u.light_mvp.m[i] = p->light_mvp[i]; u.mv.m[i] = p->mv[i]; }
    for (int i = 0; i < 9; ++i) u.nm.m[i] = p->normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.lpos[i] = p->light_position[i];
    u.si = p->shadow_intensity;
    u.kernel_order = p->kernel_order;
    u.penumbra = p->penumbra_size;
    u.max_search = p->max_search;
    u.thr = p->depth_threshold;
    // glUniform2f(shadowMapStep, 1.0/W, 1.0/H): a double quotient rounded to float
    u.stepx = (float)(1.0 / (double)p->shadow_map_width);
    u.stepy = (float)(1.0 / (double)p->shadow_map_height);
    u.pre_eval = p->pre_evaluation ? 1 : 0;
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    DepthMap D{depth_dev, p->shadow_map_width, p->shadow_map_height, (float)p->shadow_map_width,
               (float)p->shadow_map_height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    dim3 grid = grid_for(gb.w, gb.h), block(kBlockThreads);
    if (listed) {
        // the loop's pixels listed, then a lane per tap
        unsigned* count = static_cast<unsigned*>(workspace_dev);
        unsigned* list = reinterpret_cast<unsigned*>(static_cast<unsigned char*>(workspace_dev) + kHardListHeader);
        hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, s, count);
        const int side = rpcf_side(u);  // the same float loop as the kernels run
        const bool wide = side * side > 64;
        if (p->technique == GI_RPCF) {
            hipLaunchKernelGGL(k_hard_classify<GI_RPCF>, grid, block, 0, s, gb, u, out_dev, count, list);
            if (wide)
                hipLaunchKernelGGL(k_hard_taps_lanes<GI_RPCF>, dim3(expand_blocks<k_hard_taps_lanes<GI_RPCF>>()),
                                   block, 0, s, gb, D, u, out_dev, count, list);
            else
                hipLaunchKernelGGL(k_hard_taps<GI_RPCF>, dim3(expand_blocks<k_hard_taps<GI_RPCF>>()), block, 0, s,
                                   gb, D, u, out_dev, count, list);
        } else if (p->technique == GI_RBSM_FILTERED) {
            hipLaunchKernelGGL(k_hard_classify<GI_RBSM_FILTERED>, grid, block, 0, s, gb, u, out_dev, count, list);
            if (wide)
                hipLaunchKernelGGL(k_hard_taps_lanes<GI_RBSM_FILTERED>,
                                   dim3(expand_blocks<k_hard_taps_lanes<GI_RBSM_FILTERED>>()), block, 0, s, gb, D, u,
                                   out_dev, count, list);
            else
                hipLaunchKernelGGL(k_hard_taps<GI_RBSM_FILTERED>, dim3(expand_blocks<k_hard_taps<GI_RBSM_FILTERED>>()),
                                   block, 0, s, gb, D, u, out_dev, count, list);
        } else {
            // ConservativeSMSR's RPCF: a lane per tap at any loop size
            hipLaunchKernelGGL(k_hard_classify<GI_RPCF_CONSERVATIVE>, grid, block, 0, s, gb, u, out_dev, count,
                               list);
            hipLaunchKernelGGL(k_hard_taps_lanes<GI_RPCF_CONSERVATIVE>,
                               dim3(expand_blocks<k_hard_taps_lanes<GI_RPCF_CONSERVATIVE>>()), block, 0, s, gb, D, u,
                               out_dev, count, list);
        }
        return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
    }
    switch (p->technique) {
    case GI_NAIVE: hipLaunchKernelGGL(k_hard<GI_NAIVE>, grid, block, 0, s, gb, D, u, out_dev); break;
    case GI_PCF: hipLaunchKernelGGL(k_hard<GI_PCF>, grid, block, 0, s, gb, D, u, out_dev); break;
    case GI_RBSM: hipLaunchKernelGGL(k_hard<GI_RBSM>, grid, block, 0, s, gb, D, u, out_dev); break;
    default:  // GI_RBSM_CONSERVATIVE (the other techniques are listed above)
        hipLaunchKernelGGL(k_hard<GI_RBSM_CONSERVATIVE>, grid, block, 0, s, gb, D, u, out_dev);
        break;
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                        const float light_mvp[16], const float* trans_dev, const gi_soft_params* p, float* out_dev,
                        gi_stream stream) {
    SoftUniforms u;
    int rc = soft_setup(gbuffer, maps_dev, nsamples, smw, smh, light_mvp, trans_dev, p, out_dev, u);
    if (rc != GI_OK) return rc;
    GBufferView gb{gbuffer->position_dev, nullptr, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float4* tr = reinterpret_cast<const float4*>(trans_dev);
    dim3 grid(soft_blocks(gb.w, gb.h)), block(kSoftThreads);  // the soft pass's block shape
    switch (soft_variant(u)) {
    case 3: hipLaunchKernelGGL((k_soft_wide<3>), grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev); break;
    case 4: hipLaunchKernelGGL((k_soft_wide<4>), grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev); break;
    default: hipLaunchKernelGGL(k_soft<0>, grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev); break;
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_soft_shadow_pass_chunk(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                              const float light_mvp[16], const float* trans_dev, const gi_soft_params* p,
                              float* acc_dev, int first, int total_samples, float* out_dev, gi_stream stream) {
    SoftUniforms u;
    int rc = soft_setup(gbuffer, maps_dev, nsamples, smw, smh, light_mvp, trans_dev, p, out_dev, u);
    if (rc != GI_OK) return rc;
    if (!acc_dev || total_samples < 0 || total_samples > GI_MAX_SAMPLES) return GI_EINVAL;
    if (total_samples > 0 && total_samples < nsamples) return GI_EINVAL;
    SoftChunk ch{acc_dev, first ? 1 : 0, (float)total_samples};
    GBufferView gb{gbuffer->position_dev, nullptr, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float4* tr = reinterpret_cast<const float4*>(trans_dev);
    dim3 grid(soft_blocks(gb.w, gb.h)), block(kSoftThreads);  // the soft pass's block shape
    switch (soft_variant(u)) {
    case 3: hipLaunchKernelGGL(k_soft_chunk<3>, grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, ch, out_dev); break;
    case 4: hipLaunchKernelGGL(k_soft_chunk<4>, grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, ch, out_dev); break;
    default: hipLaunchKernelGGL(k_soft_chunk<0>, grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, ch, out_dev); break;
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_adaptive_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                                 const float light_mvp[16], const float* trans_dev, const gi_soft_params* p,
                                 float* out_dev, gi_stream stream) {
    SoftUniforms u;
    int rc = soft_setup(gbuffer, maps_dev, nsamples, smw, smh, light_mvp, trans_dev, p, out_dev, u);
    if (rc != GI_OK) return rc;
    GBufferView gb{gbuffer->position_dev, nullptr, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float4* tr = reinterpret_cast<const float4*>(trans_dev);
    dim3 grid(soft_blocks(gb.w, gb.h)), block(kSoftThreads);  // the soft pass's block shape
    switch (soft_variant(u)) {
    case 3:
        hipLaunchKernelGGL((k_soft_wide<3, true>), grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev);
        break;
    case 4:
        hipLaunchKernelGGL((k_soft_wide<4, true>), grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev);
        break;
    default: hipLaunchKernelGGL((k_soft<0, true>), grid, block, 0, s, gb, maps_dev, smw, smh, tr, u, out_dev); break;
    }
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_quadtree_node_test(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                          const int map_index[4], const float biased_mvps[64], int div, float* image_dev,
                          unsigned* count_dev, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !maps_dev || !map_index || !biased_mvps || !image_dev || !count_dev) return GI_EINVAL;
    if (nmaps <= 0 || smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE) return GI_EINVAL;
    if (div < 1 || div > gbuffer->width || div > gbuffer->height) return GI_EINVAL;
    for (int i = 0; i < 4; ++i)
        if (map_index[i] < 0 || map_index[i] >= nmaps) return GI_EINVAL;
    if (!finite_all(biased_mvps, 64)) return GI_ENONFINITE;
    QtUniforms u;
    for (int i = 0; i < 4; ++i) {
        for (int e = 0; e < 16; ++e) u.m[i].m[e] = biased_mvps[16 * i + e];
        u.map[i] = map_index[i];
    }
    GBufferView gb{gbuffer->position_dev, nullptr, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_qt_reproject, grid_for(gb.w, gb.h), dim3(kBlockThreads), 0, s, gb, maps_dev, smw, smh, u,
                       image_dev, count_dev);
    // glViewport(0, 0, windowWidth/div, windowHeight/div) (main.cpp:559)
    const int wd = gb.w / div, hd = gb.h / div;
    const unsigned blocks = (unsigned)(((size_t)wd * hd + 255) / 256);
    hipLaunchKernelGGL(k_qt_evaluate, dim3(blocks), dim3(256), 0, s, image_dev, gb.w, gb.h, wd, hd, count_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_rev_discontinuity_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                              int map_index, const float biased_mvp[16], const float mv[16],
                              const float normal_matrix[9], const float light_position[3], float shadow_intensity,
                              unsigned char* disc_dev, int disc_layers, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !gbuffer->normal_dev || !maps_dev || !biased_mvp || !mv || !normal_matrix ||
        !light_position || !disc_dev)
        return GI_EINVAL;
    if (nmaps <= 0 || smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE) return GI_EINVAL;
    if (map_index < 0 || map_index >= nmaps || disc_layers <= 0) return GI_EINVAL;
    if (!finite_all(biased_mvp, 16) || !finite_all(mv, 16) || !finite_all(normal_matrix, 9) ||
        !finite_all(light_position, 3) || !std::isfinite(shadow_intensity))
        return GI_ENONFINITE;
    if (map_index >= disc_layers) return GI_OK;  // no layer for this slot
    HardUniforms u{};
    for (int i = 0; i < 16; ++i) { u.light_mvp.m[i] = biased_mvp[i]; u.mv.m[i] = mv[i]; }
    for (int i = 0; i < 9; ++i) u.nm.m[i] = normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.lpos[i] = light_position[i];
    u.si = shadow_intensity;
    u.stepx = (float)(1.0 / (double)smw);
    u.stepy = (float)(1.0 / (double)smh);
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    DepthMap D{maps_dev + (size_t)map_index * smw * smh, smw, smh, (float)smw, (float)smh};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_rev_discontinuity, grid_for(gb.w, gb.h), dim3(kBlockThreads), 0, s, gb, D, u,
                       disc_dev + (size_t)map_index * gb.w * gb.h);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_rev_quadtree_node_test(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                              const int map_index[4], const float biased_mvps[64], const unsigned char* disc_dev,
                              int disc_layers, int level, int div, float* image_dev, float* visibility_dev,
                              unsigned* count_dev, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !maps_dev || !map_index || !biased_mvps || !disc_dev || !image_dev ||
        !visibility_dev || !count_dev)
        return GI_EINVAL;
    if (nmaps <= 0 || smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE) return GI_EINVAL;
    if (div < 1 || div > gbuffer->width || div > gbuffer->height || level < 0 || disc_layers <= 0)
        return GI_EINVAL;
    for (int i = 0; i < 4; ++i)
        if (map_index[i] < 0 || map_index[i] >= nmaps) return GI_EINVAL;
    if (!finite_all(biased_mvps, 64)) return GI_ENONFINITE;
    RevQtUniforms u;
    for (int i = 0; i < 4; ++i) {
        for (int e = 0; e < 16; ++e) u.q.m[i].m[e] = biased_mvps[16 * i + e];
        u.q.map[i] = map_index[i];
    }
    u.level = level;
    u.disc_layers = disc_layers;
    GBufferView gb{gbuffer->position_dev, nullptr, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_rev_qt_reproject, grid_for(gb.w, gb.h), dim3(kBlockThreads), 0, s, gb, maps_dev, smw, smh,
                       u, disc_dev, image_dev, visibility_dev, count_dev);
    const int wd = gb.w / div, hd = gb.h / div;
    const unsigned blocks = (unsigned)(((size_t)wd * hd + 255) / 256);
    hipLaunchKernelGGL(k_qt_evaluate, dim3(blocks), dim3(256), 0, s, image_dev, gb.w, gb.h, wd, hd, count_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

size_t gi_rev_soft_workspace_size(int width, int height) {
    if (width <= 0 || height <= 0 || width > GI_MAX_EDGE || height > GI_MAX_EDGE) return 0;
    return rev_soft_workspace_bytes(width, height);
}

int gi_rev_adaptive_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw,
                                     int smh, const float light_mvp[16], const float* trans_dev,
                                     const unsigned char* disc_dev, int disc_layers, const float* visibility_dev,
                                     const gi_rev_soft_params* p, float* out_dev, void* workspace_dev,
                                     size_t workspace_bytes, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !gbuffer->normal_dev || !maps_dev || !light_mvp || !trans_dev || !disc_dev ||
        !visibility_dev || !p || !out_dev || !workspace_dev)
        return GI_EINVAL;
    if ((reinterpret_cast<uintptr_t>(workspace_dev) & 15u) != 0) return GI_EINVAL;
    if (workspace_bytes < rev_soft_workspace_bytes(gbuffer->width, gbuffer->height)) return GI_ENOSPACE;
    if (nsamples <= 0 || nsamples > GI_MAX_SAMPLES) return GI_EINVAL;
    if (smw <= 0 || smh <= 0 || smw > GI_MAX_EDGE || smh > GI_MAX_EDGE || disc_layers <= 0) return GI_EINVAL;
    if (p->max_search < 0 || p->max_search > 64) return GI_EINVAL;
    if ((reinterpret_cast<uintptr_t>(trans_dev) & 15u) != 0) return GI_EINVAL;
    if (!finite_all(light_mvp, 16) || !std::isfinite(p->shadow_intensity) || !std::isfinite(p->depth_threshold) ||
        !finite_all(p->mv, 16) || !finite_all(p->normal_matrix, 9) || !finite_all(p->light_position, 3))
        return GI_ENONFINITE;
    RevSoftUniforms u{};
    for (int i = 0; i < 12; ++i) u.m[i] = light_mvp[i];
    for (int i = 0; i < 16; ++i) u.h.mv.m[i] = p->mv[i];
    for (int i = 0; i < 9; ++i) u.h.nm.m[i] = p->normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.h.lpos[i] = p->light_position[i];
    u.h.si = p->shadow_intensity;
    u.h.max_search = p->max_search;
    u.h.thr = p->depth_threshold;
    u.h.stepx = (float)(1.0 / (double)smw);
    u.h.stepy = (float)(1.0 / (double)smh);
    u.nsamples = nsamples;
    u.disc_layers = disc_layers;
    u.cap = 4 * (smw + smh) + 16;
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const float4* tr = reinterpret_cast<const float4*>(trans_dev);
    unsigned* counters = static_cast<unsigned*>(workspace_dev);
    unsigned* list = reinterpret_cast<unsigned*>(static_cast<unsigned char*>(workspace_dev) + kListHeader);
    dim3 block(kBlockThreads);
    hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL(k_rev_soft_classify, grid_for(gb.w, gb.h), block, 0, s, gb, visibility_dev, u, out_dev,
                       counters, list);
    if (p->rpcf)
        hipLaunchKernelGGL(k_rev_soft_taps, dim3(expand_blocks<k_rev_soft_taps>()), block, 0, s, gb, maps_dev, smw,
                           smh, tr, u, out_dev, counters, list);
    else
        hipLaunchKernelGGL(k_rev_soft_samples<false>, dim3(expand_blocks<k_rev_soft_samples<false>>()), block, 0, s,
                           gb, maps_dev, smw, smh, tr, disc_dev, u, out_dev, counters, list);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_phong_pass(const gi_gbuffer* gbuffer, const float* shadow_dev, const float* color_dev,
                  const gi_phong_params* p, float* rgba_out_dev, gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !gbuffer->normal_dev || !gbuffer->facing_dev || !shadow_dev || !p || !rgba_out_dev)
        return GI_EINVAL;
    if ((reinterpret_cast<uintptr_t>(rgba_out_dev) & 15u) != 0 ||
        (color_dev && (reinterpret_cast<uintptr_t>(color_dev) & 15u) != 0))
        return GI_EINVAL;
    if (!finite_all(p->mv, 16) || !finite_all(p->normal_matrix, 9) || !finite_all(p->light_position, 3) ||
        !std::isfinite(p->shadow_intensity) || !finite_all(p->color, 4) || !finite_all(p->clear_color, 4))
        return GI_ENONFINITE;
    PhongUniforms u;
    for (int i = 0; i < 16; ++i) u.mv.m[i] = p->mv[i];
    for (int i = 0; i < 9; ++i) u.nm.m[i] = p->normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.lpos[i] = p->light_position[i];
    u.si = p->shadow_intensity;
    for (int i = 0; i < 4; ++i) {
        u.color[i] = p->color[i];
        u.clear[i] = p->clear_color[i];
    }
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_phong, dim3(row_blocks(gb.w, gb.h)), dim3(kBlockThreads), 0, s, gb, gbuffer->facing_dev,
                       shadow_dev, reinterpret_cast<const float4*>(color_dev), u,
                       reinterpret_cast<float4*>(rgba_out_dev));
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

int gi_pcss_pass(const gi_gbuffer* gbuffer, const float* depth_dev, const gi_pcss_params* p, float* out_dev,
                 gi_stream stream) {
    if (!gbuffer_ok(gbuffer) || !depth_dev || !p || !out_dev) return GI_EINVAL;
    if (p->shadow_map_width <= 0 || p->shadow_map_height <= 0 || p->shadow_map_width > GI_MAX_EDGE ||
        p->shadow_map_height > GI_MAX_EDGE)
        return GI_EINVAL;
    // A size of 1 makes filterWidth 0 and every tap offset 0/0; sizes are
    // bounded so the per-pixel loops stay within 64 x 64 taps.
    if (p->blocker_search_size < 2 || p->blocker_search_size > 64 || p->kernel_size < 2 || p->kernel_size > 64)
        return GI_EINVAL;
    if (p->light_source_radius < 0 || p->light_source_radius > 4096) return GI_EINVAL;
    if (!finite_all(p->light_mvp, 16) || !std::isfinite(p->shadow_intensity)) return GI_ENONFINITE;
    if (p->pre_evaluation) {
        if (!gbuffer->normal_dev) return GI_EINVAL;
        if (!finite_all(p->mv, 16) || !finite_all(p->normal_matrix, 9) || !finite_all(p->light_position, 3))
            return GI_ENONFINITE;
    }
    HardUniforms u{};
    for (int i = 0; i < 16; ++i) { u.light_mvp.m[i] = p->light_mvp[i]; u.mv.m[i] = p->mv[i]; }
    for (int i = 0; i < 9; ++i) u.nm.m[i] = p->normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.lpos[i] = p->light_position[i];
    u.si = p->shadow_intensity;
    u.pre_eval = p->pre_evaluation ? 1 : 0;
    PcssUniforms q;
    // `if (shadowMapWidth <= 1024.0)` (:171-173): radius over the map width, or over 1024
    const float R = (float)p->light_source_radius;
    q.bsw = (p->shadow_map_width <= 1024) ? R / (float)p->shadow_map_width : R / 1024.0f;
    q.bfw = ((float)p->blocker_search_size - 1.0f) * 0.5f;
    q.kfw = ((float)p->kernel_size - 1.0f) * 0.5f;
    q.radius = R;
    q.znear = (float)p->z_near;
    q.ks = p->kernel_size;
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    DepthMap D{depth_dev, p->shadow_map_width, p->shadow_map_height, (float)p->shadow_map_width,
               (float)p->shadow_map_height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_pcss, grid_for(gb.w, gb.h), dim3(kBlockThreads), 0, s, gb, D, u, q, out_dev);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

size_t gi_rbssm_workspace_size(int width, int height) {
    if (width <= 0 || height <= 0 || width > GI_MAX_EDGE || height > GI_MAX_EDGE) return 0;
    return rbssm_workspace_bytes(width, height);
}

int gi_rbssm_pass(const gi_gbuffer* gbuffer, const float* depth_dev, const gi_rbssm_params* p, float* out_dev,
                  void* workspace_dev, size_t workspace_bytes, gi_stream stream) {
    if (!p) return GI_EINVAL;
    const gi_pcss_params& c = p->pcss;
    if (!gbuffer_ok(gbuffer) || !depth_dev || !out_dev || !workspace_dev) return GI_EINVAL;
    if ((reinterpret_cast<uintptr_t>(workspace_dev) & 15u) != 0) return GI_EINVAL;
    if (workspace_bytes < rbssm_workspace_bytes(gbuffer->width, gbuffer->height)) return GI_ENOSPACE;
    if (c.shadow_map_width <= 0 || c.shadow_map_height <= 0 || c.shadow_map_width > GI_MAX_EDGE ||
        c.shadow_map_height > GI_MAX_EDGE)
        return GI_EINVAL;
    // per-pixel work is kernel_size^2 taps x 4 walks x max_search steps: both are bounded
    if (c.blocker_search_size < 2 || c.blocker_search_size > 64 || c.kernel_size < 2 || c.kernel_size > 31)
        return GI_EINVAL;
    if (c.light_source_radius < 0 || c.light_source_radius > 4096) return GI_EINVAL;
    if (p->max_search < 0 || p->max_search > 64) return GI_EINVAL;
    if (!finite_all(c.light_mvp, 16) || !std::isfinite(c.shadow_intensity) || !std::isfinite(p->depth_threshold))
        return GI_ENONFINITE;
    if (c.pre_evaluation) {
        if (!gbuffer->normal_dev) return GI_EINVAL;
        if (!finite_all(c.mv, 16) || !finite_all(c.normal_matrix, 9) || !finite_all(c.light_position, 3))
            return GI_ENONFINITE;
    }
    HardUniforms u{};
    for (int i = 0; i < 16; ++i) { u.light_mvp.m[i] = c.light_mvp[i]; u.mv.m[i] = c.mv[i]; }
    for (int i = 0; i < 9; ++i) u.nm.m[i] = c.normal_matrix[i];
    for (int i = 0; i < 3; ++i) u.lpos[i] = c.light_position[i];
    u.si = c.shadow_intensity;
    u.pre_eval = c.pre_evaluation ? 1 : 0;
    u.max_search = p->max_search;
    u.thr = p->depth_threshold;
    // glUniform2f(shadowMapStep, 1.0/W, 1.0/H) (MyGLGeometryViewer.cpp:226-227)
    u.stepx = (float)(1.0 / (double)c.shadow_map_width);
    u.stepy = (float)(1.0 / (double)c.shadow_map_height);
    PcssUniforms q;
    const float R = (float)c.light_source_radius;
    q.bsw = (c.shadow_map_width <= 1024) ? R / (float)c.shadow_map_width : R / 1024.0f;
    q.bfw = ((float)c.blocker_search_size - 1.0f) * 0.5f;
    q.kfw = ((float)c.kernel_size - 1.0f) * 0.5f;
    q.radius = R;
    q.znear = (float)c.z_near;
    q.ks = c.kernel_size;
    GBufferView gb{gbuffer->position_dev, gbuffer->normal_dev, gbuffer->width, gbuffer->height};
    DepthMap D{depth_dev, c.shadow_map_width, c.shadow_map_height, (float)c.shadow_map_width,
               (float)c.shadow_map_height};
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    unsigned* counters = static_cast<unsigned*>(workspace_dev);
    uint2* list = reinterpret_cast<uint2*>(static_cast<unsigned char*>(workspace_dev) + kListHeader);
    hipLaunchKernelGGL(k_list_reset, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL(k_rbssm_classify, grid_for(gb.w, gb.h), dim3(kBlockThreads), 0, s, gb, D, u, q, out_dev,
                       counters, list);
    hipLaunchKernelGGL(k_rbssm_taps, dim3(expand_blocks<k_rbssm_taps>()), dim3(kBlockThreads), 0, s, gb, D, u, q,
                       out_dev, counters, list);
    return hipGetLastError() == hipSuccess ? GI_OK : GI_EDEVICE;
}

}  // extern "C"
