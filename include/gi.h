/*
 * gi.h — C ABI of the MI355X shadow pipeline (libgi.so).
 *
 * Drop-in boundary for the reference's per-technique shadow passes
 * (MarcioCerqueira/GlobalIllumination).  The reference has no FFI: its pass
 * interface is the set of `void(void)` pass functions in each app's main.cpp,
 * the `ShadowParams` struct handed by value to MyGLGeometryViewer::configure*,
 * and the GL texture contract (D32F depth texture, RGBA32F G-buffer, R = shadow
 * output).  Each entry point below names the reference function it replaces.
 *
 * Conventions
 *   - Every function returns an int status: GI_OK (0) or a negative gi_status.
 *     Nothing calls exit(); errors are never printed by the library.
 *   - Matrices are column-major float[16] (glm / OpenGL layout).
 *   - Images are row-major with row 0 at the bottom (GL window y), so texel
 *     (i, j) of a W x H map is map[j*W + i].
 *   - Pointers whose name ends in _dev are device (HBM) pointers; all other
 *     pointers are host pointers.  All buffers are owned by the caller; the
 *     passes allocate nothing and only use the caller's workspace.
 *   - `stream` is a hipStream_t (NULL = the legacy default stream).  Passes
 *     are stream-ordered and asynchronous; the library keeps no hidden state
 *     between calls, so calls on different streams are independent.
 */
#ifndef GI_H
#define GI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GI_API_VERSION 1

typedef enum gi_status {
    GI_OK = 0,
    GI_EINVAL = -1,      /* null pointer, non-positive size, bad enum, too many samples */
    GI_ENONFINITE = -2,  /* a matrix or scalar parameter is NaN / Inf */
    GI_EDEVICE = -3,     /* a HIP call failed */
    GI_ENOSPACE = -4,    /* workspace smaller than the matching *_workspace_size() */
    GI_ENOTSUP = -5,     /* size outside what the kernels support */
    GI_EIO = -6          /* a scene or OBJ file cannot be opened or parsed (gi_scene.h) */
} gi_status;

typedef void* gi_stream; /* hipStream_t */

/* Maximum area-light samples of one soft pass: AccurateSoftShadow.frag:9 caps
 * the reference at 289; HBM holds far more. */
#define GI_MAX_SAMPLES 4096
/* Largest shadow-map / view edge the raster supports (guard band arithmetic). */
#define GI_MAX_EDGE 16384

int gi_api_version(void);
const char* gi_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Light camera and texture-space projection (host-side, FP32 glm 0.9.3.1)   */
/* ------------------------------------------------------------------------ */

/* glm::perspective with fovy in DEGREES.  Replaces the projection built by
 * MyGLGeometryViewer::configureAmbient (ShadowMapping/src/Viewers/
 * MyGLGeometryViewer.cpp:14-26) via glm/gtc/matrix_transform.inl:223-244. */
int gi_perspective(float fovy_deg, float aspect, float znear, float zfar, float out[16]);
/* glm::ortho (glm/gtc/matrix_transform.inl:163-182); the reference has no
 * orthographic light, used for directional lights. */
int gi_ortho(float l, float r, float b, float t, float znear, float zfar, float out[16]);
/* glm::lookAt (glm/gtc/matrix_transform.inl:383-411), the view matrix of
 * MyGLGeometryViewer::configureAmbient. */
int gi_lookat(const float eye[3], const float center[3], const float up[3], float out[16]);
/* glm mat4 * mat4 (glm/core/type_mat4x4.inl:757-777). */
int gi_mat4_mul(const float a[16], const float b[16], float out[16]);
/* Texture-space bias matrix B of MyGLGeometryViewer::configureShadow
 * (ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:139-145).  A constant
 * window-depth bias is folded into B[3][2] = 0.5 - zbias (zbias = 0 gives the
 * reference matrix exactly). */
int gi_bias_matrix(float zbias, float out[16]);
/* UniformSampledLightSource::computeUniformSampling
 * (SoftShadowMapping/src/Scene/LightSource/UniformSampledLightSource.cpp:27-37):
 * offset `base` on a sqrt(N) x sqrt(N) grid of side `size` in world XY. */
int gi_uniform_sampling(const float base[3], int size, int nsamples, int index, float out[3]);
/* Column 3 of B*lightMVP_i for each sample (the lightMVPTrans[] uniform,
 * SoftShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:184-195).
 * light_mvps: S x 16 (already biased); trans_out: S x 4. */
int gi_soft_light_trans(const float* light_mvps, int nsamples, float* trans_out);

/* ------------------------------------------------------------------------ */
/* Light-view depth pass (a2)                                                */
/* ------------------------------------------------------------------------ */

/* Replaces renderShadowMap / displaySceneFromLightPOV
 * (ShadowMapping/src/main.cpp:221-274, :350-361) and the per-sample loop of
 * renderMonteCarlo (SoftShadowMapping/src/main.cpp:774-793): depth of every
 * triangle into an FP32 map cleared to 1.0, depth test LESS, no culling. */
typedef struct gi_raster_opts {
    float polygon_offset_factor; /* glPolygonOffset(factor, units); both 0 = off */
    float polygon_offset_units;  /* the reference apps use (4, 20), main.cpp:246 */
    void* raster_begin_event;    /* optional hipEvent_t recorded on `stream` right
                                    before the tile-raster kernel, or NULL */
    void* raster_end_event;      /* optional hipEvent_t recorded right after it */
} gi_raster_opts;

/* Workspace bytes for gi_depth_pass over `nmaps` maps of one mesh. */
size_t gi_depth_pass_workspace_size(int ntris, int nmaps, int width, int height);

/* positions_dev: nverts x 3 floats; indices_dev: ntris x 3 int32;
 * mvps_dev: nmaps x 16 floats, the unbiased P*V*M of each map (device);
 * depth_out_dev: nmaps x height x width floats. */
int gi_depth_pass(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                  const float* mvps_dev, int nmaps, int width, int height,
                  const gi_raster_opts* opts, float* depth_out_dev,
                  void* workspace_dev, size_t workspace_bytes, gi_stream stream);

/* gi_depth_pass without a workspace argument.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is a static device array of gi_depth_pass_noalloc_max_bytes()
 * (64 MiB) that each GPU the library runs on holds until the process ends;
 * the call allocates nothing.  Returns GI_ENOSPACE when
 * gi_depth_pass_workspace_size() exceeds it (e.g. 16 maps of 1024^2 with 50k
 * triangles fit; 64 maps of 2048^2 do not).  Every call on one GPU shares the
 * array: calls that overlap on a GPU (other streams, other threads) race;
 * calls ordered on one stream do not.  Other arguments as gi_depth_pass. */
size_t gi_depth_pass_noalloc_max_bytes(void);
int gi_depth_pass_noalloc(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                          const float* mvps_dev, int nmaps, int width, int height,
                          const gi_raster_opts* opts, float* depth_out_dev, gi_stream stream);

/* gi_depth_pass in two steps, so that the tile raster can run a range of
 * maps at a time: a frame can then be rastered and shaded in runs of maps
 * (gi_soft_shadow_pass_chunk) with one run's maps in memory, or with a run's
 * screen pass on another stream.  gi_depth_pass_bin bins all
 * `nmaps` maps into the workspace and rasters any map whose lists overflowed
 * it; gi_depth_pass_raster then rasters maps [map_begin, map_begin +
 * map_count) from those lists.  Both take gi_depth_pass's arguments (the
 * same workspace, untouched in between); calling raster over every map once
 * after bin gives gi_depth_pass's maps bit for bit.  opts' events, when set,
 * are recorded around the raster step's kernel.  No counterpart in the
 * reference, which rasters each map in its own GL pass. */
int gi_depth_pass_bin(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                      const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                      float* depth_out_dev, void* workspace_dev, size_t workspace_bytes, gi_stream stream);
int gi_depth_pass_raster(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                         const float* mvps_dev, int nmaps, int width, int height, const gi_raster_opts* opts,
                         int map_begin, int map_count, float* depth_out_dev, void* workspace_dev,
                         size_t workspace_bytes, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Camera G-buffer (input format of the screen passes, a6)                   */
/* ------------------------------------------------------------------------ */

/* Replaces renderGBuffer (ShadowMapping/src/main.cpp:363-372) with
 * GBuffer.vert:15-21 / GBuffer.frag:35-37: perspective-correct object-space
 * position and normal of the nearest surface.  Output layout (SoA planes):
 *   position_out_dev: 3 planes (x, y, z) of width*height floats,
 *   normal_out_dev:   3 planes, the normal NEGATED where the surface is
 *                     back-facing (gl_FrontFacing == false),
 * and zeros for background pixels (the reference clears to 0 and detects the
 * background by vertex.x == 0, Shadow.frag:244). */
size_t gi_gbuffer_pass_workspace_size(int ntris, int width, int height);
int gi_gbuffer_pass(const float* positions_dev, const float* normals_dev, int nverts,
                    const int32_t* indices_dev, int ntris, const float mvp[16],
                    int width, int height, float* position_out_dev, float* normal_out_dev,
                    float* facing_out_dev, void* workspace_dev, size_t workspace_bytes, gi_stream stream);
/* facing_out_dev (may be NULL): one plane, gl_FrontFacing as 1.0 / 0.0
 * (GBuffer.frag:36), 0 for background; the deferred Phong pass reads it to
 * recover the raw normal. */

typedef struct gi_gbuffer {
    const float* position_dev; /* 3 planes of width*height */
    const float* normal_dev;   /* 3 planes, facing-flipped; may be NULL when unused */
    int width, height;
    const float* facing_dev;   /* 1 plane, 1.0 front-facing / 0.0; may be NULL when unused */
} gi_gbuffer;

/* ------------------------------------------------------------------------ */
/* Hard-shadow screen pass (a8-a12)                                          */
/* ------------------------------------------------------------------------ */

/* Technique, i.e. the ShadowParams flags of ShadowMapping/include/Viewers/
 * ShadowParams.h:18-31 and the shader computeHardShadows picks
 * (ShadowMapping/src/main.cpp:400-414). */
typedef enum gi_technique {
    GI_NAIVE = 0,         /* Shadow.frag naive == 1 */
    GI_PCF = 1,           /* Shadow.frag bilinearPCF == 1, PCF() :86-116 */
    GI_RBSM = 2,          /* NonConservativeSMSR.frag SMSR == 1 :288-304 */
    GI_RPCF = 3,          /* NonConservativeSMSR.frag RPCFPlusSMSR == 1 :306-349 */
    GI_RBSM_FILTERED = 4, /* FilteredRBSM.frag (RSMSS) :306-349 */
    /* ConservativeSMSR.frag, bound instead of NonConservativeSMSR when
     * ShadowParams.conservative is set (ShadowMapping/src/main.cpp:406-408) */
    GI_RBSM_CONSERVATIVE = 5, /* SMSR == 1: computeShadowFromSMSR :128-144 */
    GI_RPCF_CONSERVATIVE = 6  /* RPCFPlusSMSR == 1: computeShadowFromRPCF :146-198 */
} gi_technique;

/* Mirrors ShadowParams (ShadowMapping/include/Viewers/ShadowParams.h:6-37)
 * plus the Phong uniforms configurePhong uploads for the pre-evaluation. */
typedef struct gi_shadow_params {
    float light_mvp[16];     /* B * lightMVP (configureShadow :146) */
    int   shadow_map_width;  /* shadowMapWidth */
    int   shadow_map_height; /* shadowMapHeight */
    int   technique;         /* gi_technique */
    int   kernel_order;      /* kernelOrder (PCF / RPCF loop) */
    int   penumbra_size;     /* penumbraSize (PCF half-width in texels) */
    int   max_search;        /* maxSearch (RBSM edge walk) */
    float depth_threshold;   /* depthThreshold (RBSM) */
    float shadow_intensity;  /* shadowIntensity */
    int   pre_evaluation;    /* 1: computePreEvaluationBasedOnNormalOrientation */
    float mv[16];            /* MV (camera view * model) */
    float normal_matrix[9];  /* normalMatrix, column-major mat3 */
    float light_position[3]; /* lightPosition as configurePhong uploads it */
} gi_shadow_params;

/* Device workspace of gi_hard_shadow_pass for a width x height view: a
 * 256-byte header and 4 bytes per pixel.  0 for sizes out of range. */
size_t gi_hard_shadow_workspace_size(int width, int height);

/* Replaces computeHardShadows -> displaySceneFromGBuffer(shader)
 * (ShadowMapping/src/main.cpp:400-414, :302-348).  depth_dev: one map of
 * shadow_map_height x shadow_map_width; out_dev: width*height floats holding
 * the R channel (1 = lit, shadow_intensity = shadowed, 0 = background).
 * GI_RPCF, GI_RBSM_FILTERED and GI_RPCF_CONSERVATIVE list the pixels that
 * run their tap loop in the workspace (16-byte aligned, at least
 * gi_hard_shadow_workspace_size bytes, else GI_ENOSPACE; NULL is GI_EINVAL)
 * and spread their taps over the GPU; the other techniques ignore it (NULL is
 * fine).  Calls sharing a workspace must not overlap. */
int gi_hard_shadow_pass(const gi_gbuffer* gbuffer, const float* depth_dev,
                        const gi_shadow_params* params, float* out_dev, void* workspace_dev,
                        size_t workspace_bytes, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Area-light soft-shadow screen pass (a13)                                  */
/* ------------------------------------------------------------------------ */

typedef struct gi_soft_params {
    float shadow_intensity;
    int   pcf;          /* 0: one tap per sample (AccurateSoftShadow.frag:94-97);
                           1: PCF loop per sample (:99-121) */
    int   kernel_order; /* PCF loop: step = 2*penumbra/kernel_order (reference: 3) */
    float penumbra;     /* PCF loop: half-width in texels (reference: 1) */
    int   inclusive;    /* 1: `<=` loop bound as in the reference (4x4 taps for
                           k=3); 0: `<` as Shadow.frag's PCF (3x3 taps) */
} gi_soft_params;

/* Replaces renderMonteCarlo's screen pass (SoftShadowMapping/src/main.cpp:806,
 * AccurateSoftShadow.frag:52-133).  maps_dev: nsamples x smh x smw floats;
 * light_mvp: host B*lightMVP whose columns 0-2 all samples share; trans_dev:
 * device nsamples x 4 floats (gi_soft_light_trans).  out_dev: width*height. */
int gi_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples,
                        int smw, int smh, const float light_mvp[16], const float* trans_dev,
                        const gi_soft_params* params, float* out_dev, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Adaptive area-light sampling (quad tree over the light)                    */
/* ------------------------------------------------------------------------ */

/* One quad-tree node's test in renderAdaptiveLightSourceSampling
 * (evaluateSubAreaLightSource, SoftShadowMapping/src/main.cpp:440-571):
 * QuadTree/Reprojection.frag:8-24 over the G-buffer with the node's four
 * corner maps (slots map_index[] of maps_dev, nmaps maps of smw x smh;
 * biased_mvps: B * lightMVP of each corner, 4 x 16 floats) into image_dev
 * (width*height floats: lit-corner count, 0.63 for background), then
 * QuadTree/Evaluation.frag:6-22 on the (width/div) x (height/div) viewport;
 * *count_dev receives the number of passing pixels, which the reference reads
 * back from a GL_SAMPLES_PASSED query (the node subdivides when it is > 0).
 * div: 1, or 4 with adaptiveSamplingLowerAccuracy (main.cpp:554-557). */
int gi_quadtree_node_test(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                          const int map_index[4], const float biased_mvps[64], int div, float* image_dev,
                          unsigned* count_dev, gi_stream stream);

/* gi_soft_shadow_pass over a run of consecutive samples of one frame, so that
 * a frame can be shaded in runs while later samples' maps are still being
 * rastered.  maps_dev / trans_dev point at the run's first map and trans row,
 * nsamples is the run's length; light_mvp and params as gi_soft_shadow_pass.
 * acc_dev (width*height floats, caller-owned) carries each pixel's running
 * sum from run to run: `first` starts it at 0, otherwise it is read; the
 * last run passes the frame's sample count in total_samples (> 0) and writes
 * out_dev = sum / total_samples, the earlier runs pass 0 and store the sum.
 * The additions are gi_soft_shadow_pass's in the same order, so the runs
 * give its output bit for bit.  No counterpart in the reference, whose
 * AccurateSoftShadow.frag loops over every sample in one draw. */
int gi_soft_shadow_pass_chunk(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                              const float light_mvp[16], const float* trans_dev, const gi_soft_params* params,
                              float* acc_dev, int first, int total_samples, float* out_dev, gi_stream stream);

/* The final pass of renderAdaptiveLightSourceSampling (main.cpp:868-876):
 * AccurateSoftShadow.frag with adaptiveSampling == 1.  Arguments as
 * gi_soft_shadow_pass; trans_dev[4*i+3] carries w + level * 10000 for sample
 * i (configureShadow, MyGLGeometryViewer.cpp:192), and each sample weighs
 * 1 / (2^min(level, 4) + 1)^2 in the average. */
int gi_adaptive_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw, int smh,
                                 const float light_mvp[16], const float* trans_dev, const gi_soft_params* params,
                                 float* out_dev, gi_stream stream);

/* ---- revectorization-based adaptive sampling ---------------------------- */
/* ShadowParams.revectorizationBasedAdaptiveSampling (menu entry 2,
 * SoftShadowMapping/src/main.cpp:1411-1416).  Three passes replace the plain
 * ones: every corner map rendered while the quad tree is evaluated gets a
 * screen-space discontinuity layer, the node test also keeps a per-pixel
 * visibility classification, and the final pass revectorizes the samples of
 * penumbra pixels.  The discontinuity array holds disc_layers layers of
 * width*height bytes; the reference allocates 50 (createRGBATextureArray,
 * main.cpp:1635), so maps from slot 50 on get no layer and their reads land
 * on the last layer.  Layers keep their contents from earlier frames. */
#define GI_REV_DISC_LAYERS 50

/* HardShadow/Discontinuity.frag:66-88 for corner map `map_index` (biased_mvp
 * = B * its lightMVP): per G-buffer pixel, computeDiscontinuity where the
 * pre-evaluation (mv, normal_matrix, light_position: the sample's eye as
 * configurePhong receives it during evaluation, main.cpp:365-368) and w > 0
 * let it through, else 0; background pixels keep the clear value.  Writes
 * layer map_index of disc_dev when map_index < disc_layers; otherwise nothing
 * (the reference's framebuffer is incomplete there).  The gbuffer needs its
 * normals.  Byte code: bits 0-1 r*4, bits 2-3 g*4, bit 4 discType, bit 5
 * set for an evaluated pixel, 0xFF for background. */
int gi_rev_discontinuity_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                              int map_index, const float biased_mvp[16], const float mv[16],
                              const float normal_matrix[9], const float light_position[3], float shadow_intensity,
                              unsigned char* disc_dev, int disc_layers, gi_stream stream);

/* The node test with revectorizationBasedQuadTreeEvaluation (main.cpp:
 * 506-560): QuadTree/RevectorizationBasedReprojection.frag:14-51 (lit-corner
 * count, 5 where two or more corner layers of disc_dev flag a discontinuity;
 * visibility_dev, width*height floats, set at level 0 and demoted to 0.5 at
 * deeper levels, in place), then Evaluation.frag on the (width/div) x
 * (height/div) viewport as gi_quadtree_node_test.  div follows the level:
 * 1 for levels 0-1, 2 (3 with RPCF) for 2-3, 4 (6 with RPCF) deeper
 * (main.cpp:535-549). */
int gi_rev_quadtree_node_test(const gi_gbuffer* gbuffer, const float* maps_dev, int nmaps, int smw, int smh,
                              const int map_index[4], const float biased_mvps[64], const unsigned char* disc_dev,
                              int disc_layers, int level, int div, float* image_dev, float* visibility_dev,
                              unsigned* count_dev, gi_stream stream);

typedef struct gi_rev_soft_params {
    float shadow_intensity;
    int rpcf;               /* ShadowParams.RPCF: 3 x 3 revectorized taps per sample */
    int max_search;         /* 16 (main.cpp:1617); 0..64 */
    float depth_threshold;  /* SceneLoader `d` */
    float mv[16];           /* pre-evaluation, as gi_pcss_params */
    float normal_matrix[9];
    float light_position[3];
} gi_rev_soft_params;

/* Device workspace of gi_rev_adaptive_soft_shadow_pass for a width x height
 * view: a 256-byte header and 4 bytes per pixel (the list of penumbra
 * pixels).  0 for sizes out of range. */
size_t gi_rev_soft_workspace_size(int width, int height);

/* The final pass (main.cpp:868-871): RevectorizationBasedAccurateSoftShadow.frag
 * :1252-1331.  Arguments as gi_adaptive_soft_shadow_pass plus the
 * discontinuity layers and the visibility classification of the last node
 * test; the gbuffer needs its normals.  A lit pixel takes its
 * classification clamped to [si, 1]; pixels at 0.5 average every sample,
 * revectorizing taps on a discontinuity (RSMSS :1175-1186).  The penumbra
 * pixels are listed in the workspace (16-byte aligned, at least
 * gi_rev_soft_workspace_size bytes, else GI_ENOSPACE) and their samples
 * spread over the GPU; calls sharing a workspace must not overlap. */
int gi_rev_adaptive_soft_shadow_pass(const gi_gbuffer* gbuffer, const float* maps_dev, int nsamples, int smw,
                                     int smh, const float light_mvp[16], const float* trans_dev,
                                     const unsigned char* disc_dev, int disc_layers, const float* visibility_dev,
                                     const gi_rev_soft_params* params, float* out_dev, void* workspace_dev,
                                     size_t workspace_bytes, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Percentage-closer soft shadows (plausible soft shadows, PCSS)             */
/* ------------------------------------------------------------------------ */

/* renderSoftShadows with ShadowParams.PCSS (SoftShadowMapping/src/main.cpp:
 * 925-1016): one depth map from the light centre (gi_depth_pass with
 * glPolygonOffset(4, 20), main.cpp:295), then PlausibleSoftShadow.frag:
 * blocker search (:166-192), penumbra width (:365-374) and PCF (:376-398),
 * main :605-632. */
typedef struct gi_pcss_params {
    float light_mvp[16];      /* B * lightMVP of the light centre (configureShadow :168-170) */
    int shadow_map_width, shadow_map_height;
    int blocker_search_size;  /* 7 in the reference (main.cpp:1612); 2..64 */
    int kernel_size;          /* 15 (main.cpp:1613); 2..64 */
    int light_source_radius;  /* LightSource::size / 2 (main.cpp:1614); 0..4096 */
    int z_near;               /* the int uniform zNear.  The reference never uploads it for this
                               * program (configureLinearization :88-97 runs only for the moment,
                               * exponential and filter shaders), so it stays 0 there: the
                               * penumbra width is then 0 and PCF returns 1.  Pass 0 to reproduce
                               * that, or 1 (MyGLGeometryViewer's zNear) for the intended PCSS. */
    float shadow_intensity;
    int pre_evaluation;       /* computePreEvaluationBasedOnNormalOrientation, needs normals */
    float mv[16];
    float normal_matrix[9];
    float light_position[3];
} gi_pcss_params;

/* out_dev: width*height R channel (0 for background). */
int gi_pcss_pass(const gi_gbuffer* gbuffer, const float* depth_dev, const gi_pcss_params* params, float* out_dev,
                 gi_stream stream);

/* renderSoftShadows with ShadowParams.RBSSM (SoftShadowMapping/src/main.cpp:
 * 1012): RBSSM.frag, PCSS's blocker search and penumbra width with a
 * kernelSize x kernelSize kernel whose taps on a shadow-map discontinuity are
 * revectorized (computeDiscontinuityLength :193-253, normalizeDiscontinuitySpace
 * :255-309, smoothONDS :311-1117, RBSSM :1202-1339, main :1362-1378).  The
 * pcss fields keep their meaning (kernel_size 7 in the reference's menu,
 * main.cpp:1454; z_near as documented there, 0 in the reference app). */
typedef struct gi_rbssm_params {
    gi_pcss_params pcss;
    int max_search;         /* edge walk length: 16 (main.cpp:1617); 0..64 */
    float depth_threshold;  /* SceneLoader `d` (main.cpp:1618) */
} gi_rbssm_params;

/* Device workspace of gi_rbssm_pass for a width x height view: a 256-byte
 * header and 8 bytes per pixel (the list of pixels that run the kernel
 * loop, with their penumbra widths).  0 for sizes out of range. */
size_t gi_rbssm_workspace_size(int width, int height);

/* out_dev: width*height R channel (0 for background).  kernel_size 2..31.
 * Pixels that run the kernel loop are listed in the workspace (16-byte
 * aligned, at least gi_rbssm_workspace_size bytes, else GI_ENOSPACE) and
 * their taps spread over the GPU; calls sharing a workspace must not
 * overlap. */
int gi_rbssm_pass(const gi_gbuffer* gbuffer, const float* depth_dev, const gi_rbssm_params* params, float* out_dev,
                  void* workspace_dev, size_t workspace_bytes, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Deferred Phong shading                                                     */
/* ------------------------------------------------------------------------ */

/* shadeScene / deferredShading (ShadowMapping/src/main.cpp:449-456,
 * SoftShadowMapping/src/main.cpp:1109-1117) with PhongShading.frag:11-46:
 * the shadow R channel (hard or soft pass output) modulates ambient + diffuse
 * + (shadow - si)-weighted specular of the G-buffer colour.  mv, normal_matrix
 * and light_position are configurePhong's uniforms (MyGLGeometryViewer.cpp:
 * 108-134; the light is the world light rotated 180 degrees about y, as
 * displaySceneFromGBuffer passes it, main.cpp:309). */
typedef struct gi_phong_params {
    float mv[16];
    float normal_matrix[9];
    float light_position[3];
    float shadow_intensity;
    float color[4];       /* G-buffer colour of every pixel when color_dev is NULL */
    float clear_color[4]; /* background pixels (vertex.x == 0 is discarded) */
} gi_phong_params;

/* gbuffer: position, normal and facing planes.  shadow_dev: width*height R
 * channel.  color_dev: RGBA per pixel (16-byte aligned) or NULL.
 * rgba_out_dev: RGBA per pixel (16-byte aligned). */
int gi_phong_pass(const gi_gbuffer* gbuffer, const float* shadow_dev, const float* color_dev,
                  const gi_phong_params* params, float* rgba_out_dev, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Differentiable area-light soft shadows                                     */
/* ------------------------------------------------------------------------ */

/* No reference counterpart: the reference's tests are step functions.  The
 * continuous model keeps AccurateSoftShadow.frag's structure (a depth map per
 * light sample, the mean over samples) with a smooth per-sample visibility
 *   V_i = si + (1 - si) * sum_4texels w * sigmoid((d - z + bias) / tau),
 * w = smoothstep-bilinear weights around (s*W - .5, t*H - .5), d = the plane
 * depth of the triangle covering each texel (visibility IDs held fixed).
 * Gradients are those of this model at fixed triangle visibility. */
typedef struct gi_diff_params {
    double tau;              /* sigmoid width, window-depth units */
    double bias;             /* depth bias added before the sigmoid */
    double shadow_intensity; /* si */
} gi_diff_params;

size_t gi_diff_workspace_size(int ntris, int nsamples, int smw, int smh, int view_w, int view_h);

/* mvps_dev: nsamples x 16 unbiased P*V of the samples (depth maps);
 * biased_mvps_dev: nsamples x 16 B*P*V (receivers); both FP64 column-major
 * (the geometry is evaluated in FP64; the visibility raster that fixes the
 * triangle of every texel uses their FP32 rounding).  out_dev: view pixels. */
int gi_diff_soft_forward(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                         const double* mvps_dev, const double* biased_mvps_dev, int nsamples, int smw, int smh,
                         const gi_gbuffer* gbuffer, const gi_diff_params* params, float* out_dev,
                         void* workspace_dev, size_t workspace_bytes, gi_stream stream);

/* Adjoint of gi_diff_soft_forward for dL/dout = grad_out_dev.  Recomputes the
 * visibility raster and texel depths (nothing is kept from the forward pass).
 * Writes dL/d(B*P*V_i) (receiver projection) and dL/d(P*V_i) (depth maps),
 * nsamples x 16 FP64 each, column-major; ADDS dL/dposition into
 * grad_positions_dev (nverts x 3, zeroed by the caller).  The workspace and
 * grad_positions_dev take hardware float atomics: ordinary device memory
 * (hipMalloc), not fine-grained host-coherent memory.  The vertex sums are
 * order-dependent in the last bits; the matrix gradients are reduced in a
 * fixed order. */
int gi_diff_soft_backward(const float* positions_dev, int nverts, const int32_t* indices_dev, int ntris,
                          const double* mvps_dev, const double* biased_mvps_dev, int nsamples, int smw, int smh,
                          const gi_gbuffer* gbuffer, const gi_diff_params* params, const float* grad_out_dev,
                          double* grad_biased_mvps_dev, double* grad_mvps_dev, float* grad_positions_dev,
                          void* workspace_dev, size_t workspace_bytes, gi_stream stream);

/* ------------------------------------------------------------------------ */
/* Batched frames                                                            */
/* ------------------------------------------------------------------------ */

/* No reference counterpart: the reference renders one light at a time.  A
 * batch is many independent instances, each the work of one
 * renderMonteCarlo frame (SoftShadowMapping/src/main.cpp:756-811): its light
 * samples' depth maps (gi_depth_pass) and the soft-shadow screen pass over its
 * camera G-buffer (gi_soft_shadow_pass).  Every output is bit-identical to
 * running those two passes per instance.  Instances may differ in mesh, map
 * size, sample count, view size and soft-shadow parameters. */
typedef struct gi_batch_instance {
    const float* positions_dev;  /* nverts x 3 */
    int nverts;
    const int32_t* indices_dev;  /* ntris x 3 */
    int ntris;
    const float* light_mvps_dev; /* nsamples x 16: unbiased P*V per sample (depth maps) */
    int nsamples;
    int sm_width, sm_height;
    float polygon_offset_factor, polygon_offset_units;
    float* maps_dev;             /* nsamples x sm_height x sm_width, written (16-byte aligned) */
    gi_gbuffer gbuffer;          /* camera G-buffer (position planes) */
    float light_mvp[16];         /* B*lightMVP of the last sample (columns 0-2 used) */
    const float* trans_dev;      /* nsamples x 4 (16-byte aligned) */
    gi_soft_params soft;
    float* out_dev;              /* width x height, written */
} gi_batch_instance;

/* Launch shape of a batch; gi_batch_autotune measures the candidates. */
typedef struct gi_batch_tuning {
    int tile;          /* raster tile edge: 32 or 64 */
    int raster_groups; /* persistent workgroups per CU in the raster stages: 1..8 */
    int screen_shape;  /* screen-pass pixel block: 0 = 64x4, 1 = 32x8, 2 = 16x16 (waves are row-major
                          strips), 3 = 16x16 made of four 8x8 wave tiles, 4 = 64x4 made of four 16x4
                          wave tiles */
    int screen_groups; /* persistent workgroups per CU in the screen stage: 1..8 */
} gi_batch_tuning;

typedef struct gi_batch gi_batch;             /* plan: host object, device tables in the workspace */
typedef struct gi_batch_graph gi_batch_graph; /* a plan's launches captured in a hipGraph */

enum { GI_MAX_BATCH = 4096 };

/* Device workspace a plan needs (0 when an instance is invalid). */
size_t gi_batch_workspace_size(const gi_batch_instance* inst, int n);
/* Validates every instance, lays out and uploads the plan's tables into the
 * workspace (synchronously; set-up, not frame work).  The instances' buffers
 * must stay valid while the plan is used. */
int gi_batch_create(const gi_batch_instance* inst, int n, void* workspace_dev, size_t workspace_bytes,
                    gi_batch** out);
int gi_batch_set_tuning(gi_batch* b, const gi_batch_tuning* t);
int gi_batch_get_tuning(const gi_batch* b, gi_batch_tuning* t);
/* Enqueues the whole batch: the raster stages as persistent kernels over all
 * instances, then the screen pass as one persistent kernel per PCF variant.
 * Launches only (no allocation, copy or memset), so it can be captured. */
int gi_batch_run(gi_batch* b, gi_stream stream);
/* Times every tuning candidate on `stream` (reps runs each, after one
 * warm-up), keeps the fastest in the plan and returns it with its time. */
int gi_batch_autotune(gi_batch* b, gi_stream stream, int reps, gi_batch_tuning* best, float* best_ms);
/* Captures gi_batch_run (with the plan's current tuning) into a graph. */
int gi_batch_graph_create(gi_batch* b, gi_batch_graph** out);
int gi_batch_graph_launch(gi_batch_graph* g, gi_stream stream);
void gi_batch_graph_destroy(gi_batch_graph* g);
void gi_batch_destroy(gi_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* GI_H */
