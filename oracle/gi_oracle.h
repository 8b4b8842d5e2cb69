/*
 * gi_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP32, no FMA contraction) of the reference's shadow
 * pipeline, used as the parity checker for the HIP kernels and as the "port"
 * CPU baseline.  Nothing in the product (globalillumination_amd/) links,
 * imports or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Every function cites the reference file:line it restates.  Paths are relative
 * to the reference repository root (MarcioCerqueira/GlobalIllumination).
 *
 * Parity unpinned: or_rbssm's smoothONDS fill at edge ends, corners and
 * one-texel features (RBSSM.frag:311-1117) has no closed form to check it
 * against; its straight-edge branches, edge walks and normalization are
 * pinned, the rest is a branch-by-branch transcription.
 */
#ifndef GI_ORACLE_H
#define GI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- glm 0.9.3.1 math, FP32 (a1, a4) -------------------------------------- */
/* include/glm/gtc/matrix_transform.inl:223-244 (fovy in degrees) */
void or_perspective(float fovy_deg, float aspect, float znear, float zfar, float out[16]);
/* include/glm/gtc/matrix_transform.inl:163-182 (ortho with near/far) */
void or_ortho(float l, float r, float b, float t, float znear, float zfar, float out[16]);
/* include/glm/gtc/matrix_transform.inl:383-411 */
void or_lookat(const float eye[3], const float center[3], const float up[3], float out[16]);
/* include/glm/core/type_mat4x4.inl:757-777 */
void or_mat4_mul(const float a[16], const float b[16], float out[16]);
/* ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:139-145, with the constant
 * depth bias folded into B[3][2] = 0.5 - zbias */
void or_bias_matrix(float zbias, float out[16]);
/* SoftShadowMapping/src/Scene/LightSource/UniformSampledLightSource.cpp:27-37 */
void or_uniform_sampling(const float base[3], int size, int nsamples, int index, float out[3]);

/* ---- texture sampling (a14) ------------------------------------------------ */
/* NEAREST + CLAMP_TO_BORDER(0): ShadowMapping/src/Viewers/MyGLTextureViewer.cpp:3-28 */
float or_fetch(const float* map, int w, int h, float s, float t);

/* ---- per-technique screen passes (a8-a13) ---------------------------------- */
enum { OR_NAIVE = 0, OR_PCF = 1, OR_RBSM = 2, OR_RPCF = 3, OR_RBSM_FILTERED = 4,
       OR_RBSM_CONSERVATIVE = 5, OR_RPCF_CONSERVATIVE = 6 };

typedef struct or_params {
    float light_mvp[16];       /* B * lightMVP, column-major */
    int   sm_width, sm_height;
    int   technique;
    int   kernel_order;
    int   penumbra_size;
    int   max_search;
    float depth_threshold;
    float shadow_intensity;
    int   pre_evaluation;
    float mv[16];
    float normal_matrix[9];
    float light_position[3];
} or_params;

/* pos/nrm: 3 planes of w*h floats each (normal pre-flipped by facing);
 * out: w*h floats.  Background (x == 0) writes 0. */
int or_hard_shadow(const float* pos, const float* nrm, int w, int h,
                   const float* depth, const or_params* p, float* out);

/* AccurateSoftShadow.frag:52-133.  pcf_mode 0: single tap; 1: PCF loop with
 * (kernel_order, penumbra, inclusive '<=' when inclusive != 0). */
int or_soft_shadow(const float* pos, int w, int h, const float* maps, int nsamples,
                   int smw, int smh, const float light_mvp[16], const float* trans,
                   float si, int pcf_mode, int kernel_order, float penumbra,
                   int inclusive, float* out);

/* ---- rasterization (a2, a6) ------------------------------------------------ */
/* Depth of all triangles into a w*h FP32 map cleared to 1.0, depth test LESS,
 * optional polygon offset (ShadowMapping/src/main.cpp:246-247).  Iterates each
 * triangle's bounding box. */
int or_depth_raster(const float* pos, int nv, const int* idx, int nt,
                    const float mvp[16], int w, int h, float off_factor,
                    float off_units, float* depth_out);
/* Same result computed per texel over every triangle (tiny inputs only). */
int or_depth_raster_brute(const float* pos, int nv, const int* idx, int nt,
                          const float mvp[16], int w, int h, float off_factor,
                          float off_units, float* depth_out);
/* Visibility keys ((depth bits << 32) | tri*8+sub), cleared to ~0. */
int or_visibility_raster(const float* pos, int nv, const int* idx, int nt,
                         const float mvp[16], int w, int h, uint64_t* vis_out);
/* Camera G-buffer: ShadowMapping/Shaders/GBuffer/GBuffer.vert:15-21,
 * GBuffer.frag:35-37.  pos_out / nrm_out: 3 planes each; normal pre-flipped
 * when back-facing; background = 0. */
int or_gbuffer(const float* pos, const float* nrm, int nv, const int* idx, int nt,
               const float mvp[16], int w, int h, float* pos_out, float* nrm_out, float* fac_out);

/* AccurateSoftShadow.frag with adaptiveSampling == 1 (:79-87): trans[4i+3]
 * carries w + level * 10000; samples weigh 1 / (2^min(level,4) + 1)^2. */
int or_soft_shadow_adaptive(const float* pos, int w, int h, const float* maps, int nsamples,
                            int smw, int smh, const float light_mvp[16], const float* trans,
                            float si, int pcf_mode, int kernel_order, float penumbra,
                            int inclusive, float* out);

/* One quad-tree node's test of renderAdaptiveLightSourceSampling:
 * QuadTree/Reprojection.frag into img (w*h; 0.63 for background) with the
 * maps idx[0..3] and their B*lightMVP matrices mvps (4 x 16), then
 * QuadTree/Evaluation.frag on the (w/div) x (h/div) viewport; *count = the
 * passing pixels (the reference's GL_SAMPLES_PASSED). */
int or_qt_node_test(const float* pos, int w, int h, const float* maps, int smw, int smh, const int idx[4],
                    const float mvps[64], int div, float* img, unsigned* count);

/* PlausibleSoftShadow.frag with PCSS == 1 (:166-192, :365-398, :556-563,
 * main :605-632).  pos/nrm as or_hard_shadow; p->light_mvp, mv,
 * normal_matrix, light_position, shadow_intensity, pre_evaluation,
 * sm_width/sm_height are read from `p`. */
int or_pcss(const float* pos, const float* nrm, int w, int h, const float* depth, const or_params* p,
            int blocker_search_size, int kernel_size, int light_source_radius, int z_near, float* out);

/* RBSSM.frag (main :1362-1378): PCSS's blocker search and penumbra width,
 * then the kernel_size^2 kernel with revectorized discontinuity taps.
 * p->max_search and p->depth_threshold are the edge walks' parameters. */
int or_rbssm(const float* pos, const float* nrm, int w, int h, const float* depth, const or_params* p,
             int blocker_search_size, int kernel_size, int light_source_radius, int z_near, float* out);

/* One RBSSM tap at shadow-map coordinate (x, y) and receiver depth z:
 * out[0..2] computeDiscontinuity's (r, g, b), out[3..6] the edge-walk lengths
 * left/right/down/up, out[7..8] the normalized discontinuity (x, y),
 * out[9] smoothONDS's fill before the shadow-intensity mix. */
int or_rbssm_tap(const float* depth, const or_params* p, float x, float y, float z, float* out);
/* One RBSM tap: out = {r, g, b, left, right, down, up, rel_x, rel_y, value}
 * (binary NonConservativeSMSR or, with `filtered`, FilteredRBSM's continuous
 * revectorizeShadow; value before the mix with shadowIntensity). */
int or_rbsm_tap(const float* depth, const or_params* p, int filtered, float x, float y, float z, float* out);

/* Revectorization-based adaptive sampling.  Discontinuity layers are
 * w*h bytes: bits 0-1 r*4, bits 2-3 g*4, bit 4 discType, bit 5 alpha, 0xFF
 * for background (Discontinuity.frag's RGBA8 (r, g, b, 1) and the clear
 * colour).  HardShadow/Discontinuity.frag:66-88 for one map: p->light_mvp
 * (B * its lightMVP), sm_width/sm_height, depth_threshold, max_search, and
 * the pre-evaluation's mv, normal_matrix, light_position, shadow_intensity. */
int or_rev_discontinuity(const float* pos, const float* nrm, int w, int h, const float* depth, const or_params* p,
                         unsigned char* out);

/* QuadTree/RevectorizationBasedReprojection.frag:14-51 (vis read and
 * written in place; layers clamp to disc_layers - 1) and Evaluation.frag on
 * the (w/div) x (h/div) viewport; *count = passing pixels. */
int or_rev_node_test(const float* pos, int w, int h, const float* maps, int smw, int smh, const int idx[4],
                     const float mvps[64], const unsigned char* disc, int disc_layers, int level, int div,
                     float* img, float* vis, unsigned* count);

/* RevectorizationBasedAccurateSoftShadow.frag:1252-1331 (RSMSS :1175-1186,
 * smoothONDS :320-1138 with its scans bounded by 4 (W + H) + 16 steps).
 * rpcf: ShadowParams.RPCF. */
int or_rev_soft(const float* pos, const float* nrm, int w, int h, const float* maps, int nsamples, int smw, int smh,
                const float m[16], const float* trans, const unsigned char* disc, int disc_layers, const float* vis,
                const or_params* p, int rpcf, float* out);

/* PhongShading.frag:11-46.  pos/nrm: G-buffer planes (normal flipped on back
 * faces), fac: 1.0 front / 0.0 back, shadow: w*h R channel, color: RGBA per
 * pixel or NULL (p->color everywhere); out: RGBA per pixel. */
typedef struct or_phong_params {
    float mv[16];
    float normal_matrix[9];
    float light_position[3];
    float shadow_intensity;
    float color[4];
    float clear_color[4];
} or_phong_params;

int or_phong(const float* pos, const float* nrm, const float* fac, int w, int h, const float* shadow,
             const float* color, const or_phong_params* p, float* out);

/* ---- differentiable soft shadows (continuous model, FP64) ------------------ */
/* Texel depth of the continuous model: for every texel whose visibility key
 * (or_visibility_raster) names triangle T, the window depth of T's plane
 * through its UNSNAPPED projected vertices at the texel centre; 1.0 where no
 * triangle covers the texel.  mvp: the unbiased P*V of the map (FP64). */
int or_diff_depth(const double* pos, int nv, const int* idx, int nt, const double mvp[16], int w, int h,
                  const uint64_t* vis, double* depth_out);
/* Soft visibility of the continuous model for every G-buffer pixel:
 *   S = (1/N) sum_i [si + (1 - si) * q_i],
 *   q_i = sum over the 4 texels around (s*W - .5, t*H - .5) of smoothstep-
 *         bilinear weights times sigmoid((d - z + bias) / tau),
 * (s, t, z) = perspective divide of B*M'_i * (x, 1); texels off the map have
 * depth 0.  Background pixels (x == 0) give 0.  maps: N x h x w (FP64). */
int or_diff_soft(const float* gpos, int gw, int gh, const double* maps, int nsamples, int w, int h,
                 const double* biased_mvps, double tau, double bias, double si, double* out);

/* Reverse-mode derivative of or_diff_depth + or_diff_soft (FP64) for
 * L = sum_k grad_out[k] * S[k], the triangle of every texel held fixed by
 * vis (nsamples x h x w keys of or_visibility_raster).  mvps / biased_mvps:
 * nsamples x 16 column-major (M'_i and B*M'_i).  Outputs: grad_pos (nv x 3),
 * grad_mvps = dL/dM'_i through the texel depths and grad_biased =
 * dL/d(B*M'_i) through the receivers (nsamples x 16, column-major). */
int or_diff_adjoint(const double* pos, int nv, const int* idx, int nt, const double* mvps,
                    const double* biased_mvps, const uint64_t* vis, int nsamples, int w, int h,
                    const float* gpos, int gw, int gh, double tau, double bias, double si,
                    const double* grad_out, double* grad_pos, double* grad_mvps, double* grad_biased);

/* threads used by the OpenMP loops (0 = runtime default) */
void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
