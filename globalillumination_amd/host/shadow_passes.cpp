// shadow_passes.cpp — the reference apps' pass sequence over the C ABI.
#include "shadow_passes.hpp"

#include <cmath>
#include <memory>
#include <initializer_list>
#include <cstring>

namespace gi::host {

namespace {

#define GI_TRY(expr)                         \
    do {                                     \
        int gi_rc_ = (expr);                 \
        if (gi_rc_ != GI_OK) return gi_rc_;  \
    } while (0)

int hip_ok(hipError_t e) { return e == hipSuccess ? GI_OK : GI_EDEVICE; }

Mat4 identity() {
    Mat4 m{};
    m[0] = m[5] = m[10] = m[15] = 1.0f;
    return m;
}

Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r{};
    gi_mat4_mul(a.data(), b.data(), r.data());
    return r;
}

int upload(DeviceBuffer& buf, const void* src, size_t bytes, hipStream_t s) {
    GI_TRY(buf.allocate(bytes));
    if (bytes == 0) return GI_OK;
    GI_TRY(hip_ok(hipMemcpyAsync(buf.get(), src, bytes, hipMemcpyHostToDevice, s)));
    return hip_ok(hipStreamSynchronize(s));
}

// glm::inverseTranspose(mat3(mv)) as configurePhong uploads it
// (ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:113-116); it only feeds
// the sign test of the pre-evaluation, so it is formed in double and rounded.
void normal_matrix(const Mat4& mv, float out[9]) {
    double m[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) m[c][r] = mv[c * 4 + r];
    double cof[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            int c1 = (c + 1) % 3, c2 = (c + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
            cof[c][r] = m[c1][r1] * m[c2][r2] - m[c1][r2] * m[c2][r1];
        }
    double det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2];
    // inverse(M)^T = cofactor(M) / det
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out[c * 3 + r] = static_cast<float>(cof[c][r] / det);
}

size_t bytes_of(int n, size_t each) { return static_cast<size_t>(n) * each; }

}  // namespace

DeviceBuffer::~DeviceBuffer() {
    if (ptr_) (void)hipFree(ptr_);
}

int DeviceBuffer::allocate(size_t bytes) {
    if (bytes <= bytes_ && ptr_) return GI_OK;
    if (ptr_) {
        (void)hipFree(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    if (bytes == 0) return GI_OK;
    if (hipMalloc(&ptr_, bytes) != hipSuccess) {
        ptr_ = nullptr;
        return GI_EDEVICE;
    }
    bytes_ = bytes;
    return GI_OK;
}

int GeometryViewer::configureAmbient(int width, int height) {
    GI_TRY(gi_perspective(fov, static_cast<float>(width) / static_cast<float>(height), zNear, zFar,
                          projection.data()));
    GI_TRY(gi_lookat(eye.data(), look.data(), up.data(), view.data()));
    model = identity();
    return GI_OK;
}

Mat4 GeometryViewer::mvp() const { return mul(mul(projection, view), model); }

// ---------------------------------------------------------------------------
// HardShadowPipeline
// ---------------------------------------------------------------------------
HardShadowPipeline::~HardShadowPipeline() {
    if (ownStream_ && stream) (void)hipStreamDestroy(stream);
}

int HardShadowPipeline::init(const Mesh& mesh, int windowWidth, int windowHeight, int shadowMapWidth,
                             int shadowMapHeight) {
    if (windowWidth <= 0 || windowHeight <= 0 || shadowMapWidth <= 0 || shadowMapHeight <= 0) return GI_EINVAL;
    if (mesh.nverts() <= 0 || mesh.normals.size() != mesh.positions.size()) return GI_EINVAL;
    if (!stream) {
        GI_TRY(hip_ok(hipStreamCreate(&stream)));
        ownStream_ = true;
    }
    windowWidth_ = windowWidth;
    windowHeight_ = windowHeight;
    nverts_ = mesh.nverts();
    ntris_ = mesh.ntris();
    shadowParams.shadowMapWidth = shadowMapWidth;
    shadowParams.shadowMapHeight = shadowMapHeight;
    GI_TRY(upload(pos_, mesh.positions.data(), mesh.positions.size() * sizeof(float), stream));
    GI_TRY(upload(nrm_, mesh.normals.data(), mesh.normals.size() * sizeof(float), stream));
    GI_TRY(upload(idx_, mesh.indices.data(), mesh.indices.size() * sizeof(int32_t), stream));
    GI_TRY(mvp_.allocate(16 * sizeof(float)));
    GI_TRY(depth_.allocate(bytes_of(shadowMapWidth * shadowMapHeight, sizeof(float))));
    const int npx = windowWidth * windowHeight;
    GI_TRY(gpos_.allocate(bytes_of(npx, 3 * sizeof(float))));
    GI_TRY(gnrm_.allocate(bytes_of(npx, 3 * sizeof(float))));
    GI_TRY(gfac_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(hard_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(rgba_.allocate(bytes_of(npx, 4 * sizeof(float))));
    // one workspace for the frame's passes, which run one after another
    size_t need = 0;
    for (size_t b : {gi_depth_pass_workspace_size(ntris_, 1, shadowMapWidth, shadowMapHeight),
                     gi_gbuffer_pass_workspace_size(ntris_, windowWidth, windowHeight),
                     gi_hard_shadow_workspace_size(windowWidth, windowHeight)}) {
        if (b == 0) return GI_ENOTSUP;
        need = b > need ? b : need;
    }
    return ws_.allocate(need);
}

// displaySceneFromLightPOV (main.cpp:221-274) inside renderShadowMap
// (:350-361): lightMVP = P*V*M of the light, depth of every triangle.
int HardShadowPipeline::renderShadowMap() {
    GeometryViewer light;
    light.eye = lightEye;
    light.look = lightAt;
    light.up = lightUp;
    GI_TRY(light.configureAmbient(shadowParams.shadowMapWidth, shadowParams.shadowMapHeight));
    shadowParams.lightMVP = light.mvp();
    shadowParams.lightMV = mul(light.view, light.model);
    shadowParams.lightP = light.projection;
    GI_TRY(hip_ok(hipMemcpyAsync(mvp_.get(), shadowParams.lightMVP.data(), 16 * sizeof(float),
                                 hipMemcpyHostToDevice, stream)));
    return gi_depth_pass(pos_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_, mvp_.as<float>(), 1,
                         shadowParams.shadowMapWidth, shadowParams.shadowMapHeight, &rasterOpts,
                         depth_.as<float>(), ws_.get(), ws_.size(), stream);
}

// displaySceneFromCameraPOV(GBUFFER_SHADER) (main.cpp:276-300, :363-372)
int HardShadowPipeline::renderGBuffer() {
    GeometryViewer cam;
    cam.fov = cameraFov;
    cam.eye = cameraEye;
    cam.look = cameraAt;
    cam.up = cameraUp;
    GI_TRY(cam.configureAmbient(windowWidth_, windowHeight_));
    cameraMVP_ = cam.mvp();
    cameraMV_ = mul(cam.view, cam.model);
    return gi_gbuffer_pass(pos_.as<float>(), nrm_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_,
                           cameraMVP_.data(), windowWidth_, windowHeight_, gpos_.as<float>(), gnrm_.as<float>(),
                           gfac_.as<float>(), ws_.get(), ws_.size(), stream);
}

// The shader computeHardShadows picks from the flags (main.cpp:400-414).
int HardShadowPipeline::technique() const {
    const ShadowParams& p = shadowParams;
    if (p.SMSR) return p.conservative ? GI_RBSM_CONSERVATIVE : GI_RBSM;
    if (p.RPCFPlusSMSR) return p.conservative ? GI_RPCF_CONSERVATIVE : GI_RPCF;
    if (p.RSMSS || p.RPCFPlusRSMSS) return GI_RBSM_FILTERED;
    if (p.naive) return GI_NAIVE;
    return GI_PCF;  // Shadow.frag falls through to PCF (:265-266)
}

int HardShadowPipeline::computeHardShadows() {
    const ShadowParams& sp = shadowParams;
    if (!sp.naive && !sp.SMSR && !sp.RPCFPlusSMSR && !sp.RSMSS && !sp.RPCFPlusRSMSS && !sp.bilinearPCF)
        return GI_ENOTSUP;  // Shadow.frag's PCF reads an unset depth without bilinearPCF (:101-104)
    gi_shadow_params p{};
    float B[16];
    GI_TRY(gi_bias_matrix(sp.zBias, B));
    GI_TRY(gi_mat4_mul(B, sp.lightMVP.data(), p.light_mvp));  // configureShadow :139-146
    p.shadow_map_width = sp.shadowMapWidth;
    p.shadow_map_height = sp.shadowMapHeight;
    p.technique = technique();
    p.kernel_order = sp.kernelOrder;
    p.penumbra_size = sp.penumbraSize;
    p.max_search = sp.maxSearch;
    p.depth_threshold = sp.depthThreshold;
    p.shadow_intensity = sp.shadowIntensity;
    p.pre_evaluation = sp.preEvaluation ? 1 : 0;
    // configurePhong(lightEye rotated 180 degrees about y, cameraEye) (main.cpp:309)
    std::memcpy(p.mv, cameraMV_.data(), sizeof(p.mv));
    normal_matrix(cameraMV_, p.normal_matrix);
    p.light_position[0] = -lightEye[0];
    p.light_position[1] = lightEye[1];
    p.light_position[2] = -lightEye[2];
    gi_gbuffer g{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
    return gi_hard_shadow_pass(&g, depth_.as<float>(), &p, hard_.as<float>(), ws_.get(), ws_.size(), stream);
}

// configurePhong(lightEye rotated 180 degrees about y, cameraEye)
// (MyGLGeometryViewer.cpp:108-134, main.cpp:309) with the shadow buffer a
// screen pass wrote.
static int phong(const gi_gbuffer& g, const Mat4& cameraMV, const Vec3& lightEye, float shadowIntensity,
                 const float* shadow, float* rgba, hipStream_t stream) {
    gi_phong_params p{};
    std::memcpy(p.mv, cameraMV.data(), sizeof(p.mv));
    normal_matrix(cameraMV, p.normal_matrix);
    p.light_position[0] = -lightEye[0];
    p.light_position[1] = lightEye[1];
    p.light_position[2] = -lightEye[2];
    p.shadow_intensity = shadowIntensity;
    std::memcpy(p.color, kSceneColor, sizeof(p.color));
    std::memcpy(p.clear_color, kClearColor, sizeof(p.clear_color));
    return gi_phong_pass(&g, shadow, nullptr, &p, rgba, stream);
}

std::array<float, 9> HardShadowPipeline::normalMatrix() const {
    std::array<float, 9> nm{};
    normal_matrix(cameraMV_, nm.data());
    return nm;
}

int HardShadowPipeline::shadeScene() {
    gi_gbuffer g{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
    return phong(g, cameraMV_, lightEye, shadowParams.shadowIntensity, hard_.as<float>(), rgba_.as<float>(), stream);
}

int HardShadowPipeline::display() {
    GI_TRY(renderShadowMap());
    GI_TRY(renderGBuffer());
    GI_TRY(computeHardShadows());
    return shadeScene();
}

int HardShadowPipeline::download(std::vector<float>& hardShadow, std::vector<float>* depth,
                                 std::vector<float>* gbufferPosition, std::vector<float>* gbufferNormal,
                                 std::vector<float>* rgba) {
    const size_t npx = static_cast<size_t>(windowWidth_) * windowHeight_;
    hardShadow.resize(npx);
    GI_TRY(hip_ok(hipMemcpyAsync(hardShadow.data(), hard_.get(), npx * sizeof(float), hipMemcpyDeviceToHost,
                                 stream)));
    if (depth) {
        depth->resize(static_cast<size_t>(shadowParams.shadowMapWidth) * shadowParams.shadowMapHeight);
        GI_TRY(hip_ok(hipMemcpyAsync(depth->data(), depth_.get(), depth->size() * sizeof(float),
                                     hipMemcpyDeviceToHost, stream)));
    }
    if (gbufferPosition) {
        gbufferPosition->resize(3 * npx);
        GI_TRY(hip_ok(hipMemcpyAsync(gbufferPosition->data(), gpos_.get(), 3 * npx * sizeof(float),
                                     hipMemcpyDeviceToHost, stream)));
    }
    if (gbufferNormal) {
        gbufferNormal->resize(3 * npx);
        GI_TRY(hip_ok(hipMemcpyAsync(gbufferNormal->data(), gnrm_.get(), 3 * npx * sizeof(float),
                                     hipMemcpyDeviceToHost, stream)));
    }
    if (rgba) {
        rgba->resize(4 * npx);
        GI_TRY(hip_ok(hipMemcpyAsync(rgba->data(), rgba_.get(), 4 * npx * sizeof(float), hipMemcpyDeviceToHost,
                                     stream)));
    }
    return hip_ok(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------
// SoftShadowPipeline
// ---------------------------------------------------------------------------
SoftShadowPipeline::~SoftShadowPipeline() {
    if (ownStream_ && stream) (void)hipStreamDestroy(stream);
}

int SoftShadowPipeline::init(const Mesh& mesh, int windowWidth, int windowHeight, int shadowMapWidth,
                             int shadowMapHeight, int maxSamples) {
    if (windowWidth <= 0 || windowHeight <= 0 || shadowMapWidth <= 0 || shadowMapHeight <= 0) return GI_EINVAL;
    if (maxSamples <= 0 || maxSamples > GI_MAX_SAMPLES) return GI_EINVAL;
    if (mesh.nverts() <= 0 || mesh.normals.size() != mesh.positions.size()) return GI_EINVAL;
    if (!stream) {
        GI_TRY(hip_ok(hipStreamCreate(&stream)));
        ownStream_ = true;
    }
    windowWidth_ = windowWidth;
    windowHeight_ = windowHeight;
    smw_ = shadowMapWidth;
    smh_ = shadowMapHeight;
    maxSamples_ = maxSamples;
    nverts_ = mesh.nverts();
    ntris_ = mesh.ntris();
    GI_TRY(upload(pos_, mesh.positions.data(), mesh.positions.size() * sizeof(float), stream));
    GI_TRY(upload(nrm_, mesh.normals.data(), mesh.normals.size() * sizeof(float), stream));
    GI_TRY(upload(idx_, mesh.indices.data(), mesh.indices.size() * sizeof(int32_t), stream));
    GI_TRY(mvps_.allocate(bytes_of(maxSamples, 16 * sizeof(float))));
    GI_TRY(trans_.allocate(bytes_of(maxSamples, 4 * sizeof(float))));
    GI_TRY(maps_.allocate(static_cast<size_t>(maxSamples) * smw_ * smh_ * sizeof(float)));
    const int npx = windowWidth * windowHeight;
    GI_TRY(gpos_.allocate(bytes_of(npx, 3 * sizeof(float))));
    GI_TRY(gnrm_.allocate(bytes_of(npx, 3 * sizeof(float))));
    GI_TRY(gfac_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(soft_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(rgba_.allocate(bytes_of(npx, 4 * sizeof(float))));
    GI_TRY(qtImg_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(qtCount_.allocate(256));
    // revectorization-based adaptive sampling: the visibility map and the
    // 50-layer discontinuity array (createRGBATextureArray, main.cpp:1635),
    // whose layers persist across frames; they start at zero here
    GI_TRY(qtVis_.allocate(bytes_of(npx, sizeof(float))));
    GI_TRY(disc_.allocate(static_cast<size_t>(GI_REV_DISC_LAYERS) * npx));
    GI_TRY(hip_ok(hipMemsetAsync(disc_.get(), 0, disc_.size(), stream)));
    GI_TRY(hip_ok(hipMemsetAsync(qtVis_.get(), 0, qtVis_.size(), stream)));
    // one workspace for the passes of a frame, which run one after another
    size_t need = 0;
    for (size_t b : {gi_depth_pass_workspace_size(ntris_, maxSamples, smw_, smh_),
                     gi_gbuffer_pass_workspace_size(ntris_, windowWidth, windowHeight),
                     gi_rbssm_workspace_size(windowWidth, windowHeight),
                     gi_rev_soft_workspace_size(windowWidth, windowHeight)}) {
        if (b == 0) return GI_ENOTSUP;
        need = b > need ? b : need;
    }
    return ws_.allocate(need);
}

int SoftShadowPipeline::renderGBuffer() {
    GeometryViewer cam;
    cam.fov = cameraFov;
    cam.eye = cameraEye;
    cam.look = cameraAt;
    cam.up = cameraUp;
    GI_TRY(cam.configureAmbient(windowWidth_, windowHeight_));
    cameraMVP_ = cam.mvp();
    cameraMV_ = mul(cam.view, cam.model);
    return gi_gbuffer_pass(pos_.as<float>(), nrm_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_,
                           cameraMVP_.data(), windowWidth_, windowHeight_, gpos_.as<float>(), gnrm_.as<float>(),
                           gfac_.as<float>(), ws_.get(), ws_.size(), stream);
}

int SoftShadowPipeline::renderMonteCarlo() {
    const int n = numberOfSamples;
    if (n <= 1 || n > maxSamples_) return GI_EINVAL;
    lightMVPs_.assign(n, Mat4{});
    std::vector<float> biased(static_cast<size_t>(n) * 16), trans(static_cast<size_t>(n) * 4);
    float B[16];
    GI_TRY(gi_bias_matrix(0.0f, B));
    for (int i = 0; i < n; ++i) {
        // UniformSampledLightSource::getEye / getAt (:13-25): both shifted alike
        GeometryViewer light;
        GI_TRY(gi_uniform_sampling(lightEye.data(), lightSize, n, i, light.eye.data()));
        GI_TRY(gi_uniform_sampling(lightAt.data(), lightSize, n, i, light.look.data()));
        light.up = lightUp;
        GI_TRY(light.configureAmbient(smw_, smh_));
        lightMVPs_[i] = light.mvp();  // shadowParams.lightMVPs[i] = lightMVP (:790)
        GI_TRY(gi_mat4_mul(B, lightMVPs_[i].data(), &biased[16 * i]));
    }
    GI_TRY(gi_soft_light_trans(biased.data(), n, trans.data()));  // configureShadow :184-195
    GI_TRY(hip_ok(hipMemcpyAsync(mvps_.get(), lightMVPs_.data(), bytes_of(n, 16 * sizeof(float)),
                                 hipMemcpyHostToDevice, stream)));
    GI_TRY(hip_ok(hipMemcpyAsync(trans_.get(), trans.data(), bytes_of(n, 4 * sizeof(float)),
                                 hipMemcpyHostToDevice, stream)));
    gi_raster_opts opts{4.0f, 20.0f, nullptr, nullptr};  // glPolygonOffset(4, 20), Soft main.cpp:295
    GI_TRY(gi_depth_pass(pos_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_, mvps_.as<float>(), n, smw_, smh_,
                         &opts, maps_.as<float>(), ws_.get(), ws_.size(), stream));
    gi_gbuffer g{gpos_.as<float>(), nullptr, windowWidth_, windowHeight_, nullptr};
    gi_soft_params sp{shadowIntensity, adaptiveSamplingLowerAccuracy ? 1 : 0, 3, 1.0f, 1};
    // columns 0-2 come from the last sample's B*lightMVP (MyGLGeometryViewer.cpp:168-170)
    GI_TRY(gi_soft_shadow_pass(&g, maps_.as<float>(), n, smw_, smh_, &biased[16 * (n - 1)], trans_.as<float>(),
                               &sp, soft_.as<float>(), stream));
    return hip_ok(hipStreamSynchronize(stream));  // the trans upload's host buffer goes out of scope
}

int SoftShadowPipeline::renderSoftShadows() {
    // displaySceneFromLightPOV (:274-318): the light source's own view
    GeometryViewer light;
    light.eye = lightEye;
    light.look = lightAt;
    light.up = lightUp;
    GI_TRY(light.configureAmbient(smw_, smh_));
    lightMVP_ = light.mvp();
    GI_TRY(hip_ok(hipMemcpyAsync(mvps_.get(), lightMVP_.data(), 16 * sizeof(float), hipMemcpyHostToDevice,
                                 stream)));
    gi_raster_opts opts{4.0f, 20.0f, nullptr, nullptr};  // glPolygonOffset(4, 20), :295
    GI_TRY(gi_depth_pass(pos_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_, mvps_.as<float>(), 1, smw_, smh_,
                         &opts, maps_.as<float>(), ws_.get(), ws_.size(), stream));
    gi_pcss_params p{};
    float B[16];
    GI_TRY(gi_bias_matrix(0.0f, B));
    GI_TRY(gi_mat4_mul(B, lightMVP_.data(), p.light_mvp));  // configureShadow :168-170
    p.shadow_map_width = smw_;
    p.shadow_map_height = smh_;
    p.blocker_search_size = blockerSearchSize;
    p.kernel_size = kernelSize;
    p.light_source_radius = lightSize / 2;  // main.cpp:1614 (int division)
    p.z_near = pcssZNear;
    p.shadow_intensity = shadowIntensity;
    p.pre_evaluation = preEvaluation ? 1 : 0;
    // configurePhong(lightEye rotated 180 degrees about y) (displaySceneFromGBuffer :366-367, :418)
    std::memcpy(p.mv, cameraMV_.data(), sizeof(p.mv));
    normal_matrix(cameraMV_, p.normal_matrix);
    p.light_position[0] = -lightEye[0];
    p.light_position[1] = lightEye[1];
    p.light_position[2] = -lightEye[2];
    gi_gbuffer g{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
    if (rbssm) {
        gi_rbssm_params r{p, maxSearch, depthThreshold};
        return gi_rbssm_pass(&g, maps_.as<float>(), &r, soft_.as<float>(), ws_.get(), ws_.size(), stream);
    }
    return gi_pcss_pass(&g, maps_.as<float>(), &p, soft_.as<float>(), stream);
}

// ---------------------------------------------------------------------------
// Adaptive light sampling (renderAdaptiveLightSourceSampling, main.cpp:813-892)
// ---------------------------------------------------------------------------

// QuadTreeLightSource (src/Scene/LightSource/QuadTreeLightSource.cpp): a node
// holds the centre ("reference sample") of its square of the light; its four
// corner samples and its children's centres are offsets of that centre.
struct SoftShadowPipeline::QuadNode {
    Vec3 refEye{}, refAt{};
    int level = 0;
    bool hasChildren = false;
    std::unique_ptr<QuadNode> nodes[4];

    // computeUniformSamples (levelShift 1) / computeReferenceSamples (levelShift 2)
    static Vec3 offset(Vec3 v, int level, int size, int levelShift, int index) {
        const int div = static_cast<int>(std::pow(2.0f, static_cast<float>(level + levelShift)));
        const float step = static_cast<float>(size) / static_cast<float>(div);
        v[0] = v[0] + (-step + (2.0f * step) * static_cast<float>(index % 2));
        v[1] = v[1] + (-step + (2.0f * step) * static_cast<float>(index / 2));
        return v;
    }
    void subdivide(int size) {  // QuadTreeLightSource::subdivide
        for (int n = 0; n < 4; ++n) {
            nodes[n] = std::make_unique<QuadNode>();
            nodes[n]->refEye = offset(refEye, level, size, 2, n);
            nodes[n]->refAt = offset(refAt, level, size, 2, n);
            nodes[n]->level = level + 1;
        }
        hasChildren = true;
    }
};

namespace {

// computeChildIndices (main.cpp:426-438): node indices are row-major in the
// level's 2^level x 2^level grid.
void child_indices(int level, int nodeIndex, int child[4]) {
    const int div = 1 << level;
    const int x = nodeIndex % div, y = nodeIndex / div;
    const int base = 2 * x + (1 << (level + 2)) * y;
    child[0] = base;
    child[1] = base + 1;
    child[2] = base + (1 << (level + 1));
    child[3] = base + (1 << (level + 1)) + 1;
}

// The node's corner `index` on the 17 x 17 grid of samples (main.cpp:447-448).
void corner(int level, int nodeIndex, int index, int& x, int& y) {
    const int childFactor = 16 >> level, parentFactor = 1 << level;
    x = (index % 2) * childFactor + (nodeIndex % parentFactor) * childFactor;
    y = (index / 2) * childFactor + (nodeIndex / parentFactor) * childFactor;
}

}  // namespace

// displaySceneFromLightPOV for one light sample, into the next map slot.
int SoftShadowPipeline::renderLightSample(const Vec3& eye, const Vec3& at) {
    if (adaptiveSamples_ >= maxSamples_) return GI_ENOSPACE;
    GeometryViewer light;
    light.eye = eye;
    light.look = at;
    light.up = lightUp;
    GI_TRY(light.configureAmbient(smw_, smh_));
    lightMVP_ = light.mvp();
    const int n = adaptiveSamples_;
    lightMVPs_.push_back(lightMVP_);
    GI_TRY(hip_ok(hipMemcpyAsync(mvps_.as<float>() + 16 * (size_t)n, lightMVP_.data(), 16 * sizeof(float),
                                 hipMemcpyHostToDevice, stream)));
    gi_raster_opts opts{4.0f, 20.0f, nullptr, nullptr};  // glPolygonOffset(4, 20), :295
    GI_TRY(gi_depth_pass(pos_.as<float>(), nverts_, idx_.as<int32_t>(), ntris_, mvps_.as<float>() + 16 * (size_t)n,
                         1, smw_, smh_, &opts, maps_.as<float>() + (size_t)n * smw_ * smh_, ws_.get(), ws_.size(),
                         stream));
    accFactor_.push_back(0.0f);
    ++adaptiveSamples_;
    return GI_OK;
}

// The evaluation grid divisor of a node test (main.cpp:533-557).
static int node_test_div(bool revectorized, bool rpcf, bool lowerAccuracy, int level) {
    if (revectorized) {
        if (level <= 1) return 1;
        if (level <= 3) return rpcf ? 3 : 2;
        return rpcf ? 6 : 4;
    }
    return lowerAccuracy ? 4 : 1;
}

// evaluateSubAreaLightSource (main.cpp:440-571).
int SoftShadowPipeline::evaluateSubAreaLightSource(const QuadNode& node, int nodeIndex, bool& subdivide) {
    int local[4];
    float B[16], biased[64];
    GI_TRY(gi_bias_matrix(0.0f, B));
    for (int index = 0; index < 4; ++index) {
        int x, y;
        corner(node.level, nodeIndex, index, x, y);
        if (qtRendered_[x][y]) {
            local[index] = qtHash_[x][y];
            continue;
        }
        qtRendered_[x][y] = true;
        qtHash_[x][y] = adaptiveSamples_;
        local[index] = adaptiveSamples_;
        const Vec3 eye = QuadNode::offset(node.refEye, node.level, lightSize, 1, index);
        GI_TRY(renderLightSample(eye, QuadNode::offset(node.refAt, node.level, lightSize, 1, index)));
        if (revectorizationBasedAdaptiveSampling) {
            // Discontinuity.frag into layer `slot` (:476-491); configurePhong
            // receives the sample's eye as it is (no 180-degree turn while
            // evaluating, :365-368)
            float mvp[16], nm[9];
            GI_TRY(gi_mat4_mul(B, lightMVP_.data(), mvp));
            normal_matrix(cameraMV_, nm);
            gi_gbuffer gn{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
            GI_TRY(gi_rev_discontinuity_pass(&gn, maps_.as<float>(), adaptiveSamples_, smw_, smh_,
                                             local[index], mvp, cameraMV_.data(), nm, eye.data(), shadowIntensity,
                                             disc_.as<unsigned char>(), GI_REV_DISC_LAYERS, stream));
        }
    }
    // configureShadow with quadTreeEvaluation: B * lightMVPs[hash] (MyGLGeometryViewer.cpp:173-183)
    for (int i = 0; i < 4; ++i) GI_TRY(gi_mat4_mul(B, lightMVPs_[local[i]].data(), biased + 16 * i));
    gi_gbuffer g{gpos_.as<float>(), nullptr, windowWidth_, windowHeight_, nullptr};
    const int div = node_test_div(revectorizationBasedAdaptiveSampling, RPCF, adaptiveSamplingLowerAccuracy,
                                  node.level);
    if (revectorizationBasedAdaptiveSampling)
        GI_TRY(gi_rev_quadtree_node_test(&g, maps_.as<float>(), adaptiveSamples_, smw_, smh_, local, biased,
                                         disc_.as<unsigned char>(), GI_REV_DISC_LAYERS, node.level, div,
                                         qtImg_.as<float>(), qtVis_.as<float>(), qtCount_.as<unsigned>(), stream));
    else
        GI_TRY(gi_quadtree_node_test(&g, maps_.as<float>(), adaptiveSamples_, smw_, smh_, local, biased, div,
                                     qtImg_.as<float>(), qtCount_.as<unsigned>(), stream));
    unsigned count = 0;
    GI_TRY(hip_ok(hipMemcpyAsync(&count, qtCount_.get(), sizeof(count), hipMemcpyDeviceToHost, stream)));
    GI_TRY(hip_ok(hipStreamSynchronize(stream)));  // the occlusion query's GL_QUERY_RESULT
    nodeTests_.push_back({nodeIndex, node.level, count});
    subdivide = count > 0;
    return GI_OK;
}

// evaluateAreaLightSource (main.cpp:573-620), refinement only.
int SoftShadowPipeline::evaluateAreaLightSource(QuadNode& node, int nodeIndex) {
    if (node.level == maxLevel) return GI_OK;
    int child[4];
    child_indices(node.level, nodeIndex, child);
    bool subdivide = false;
    GI_TRY(evaluateSubAreaLightSource(node, nodeIndex, subdivide));
    if (subdivide) {
        node.subdivide(lightSize);
        for (int c = 0; c < 4; ++c) GI_TRY(evaluateAreaLightSource(*node.nodes[c], child[c]));
    }
    return GI_OK;
}

// renderAreaLightSource / renderSubAreaLightSource (main.cpp:622-694): each
// leaf's corners; a map already rendered takes the leaf's level.
int SoftShadowPipeline::renderAreaLightSource(const QuadNode& node, int nodeIndex) {
    if (node.hasChildren) {
        int child[4];
        child_indices(node.level, nodeIndex, child);
        for (int c = 0; c < 4; ++c) GI_TRY(renderAreaLightSource(*node.nodes[c], child[c]));
        return GI_OK;
    }
    for (int sample = 0; sample < 4; ++sample) {
        int x, y;
        corner(node.level, nodeIndex, sample, x, y);
        if (qtRendered_[x][y]) {
            accFactor_[(size_t)qtHash_[x][y]] = static_cast<float>(node.level);
            continue;
        }
        qtRendered_[x][y] = true;
        qtHash_[x][y] = adaptiveSamples_;
        GI_TRY(renderLightSample(QuadNode::offset(node.refEye, node.level, lightSize, 1, sample),
                                 QuadNode::offset(node.refAt, node.level, lightSize, 1, sample)));
        accFactor_.back() = static_cast<float>(node.level);
    }
    return GI_OK;
}

int SoftShadowPipeline::renderAdaptiveLightSourceSampling() {
    if (maxLevel < 0 || maxLevel > 4) return GI_EINVAL;  // the 17 x 17 sample grid
    for (auto& row : qtRendered_)
        for (bool& r : row) r = false;
    for (auto& row : qtHash_)
        for (int& h : row) h = -1;
    lightMVPs_.clear();
    nodeTests_.clear();
    accFactor_.clear();
    adaptiveSamples_ = 0;
    QuadNode root;  // buildFirstLevel: level 0, centred on the light
    root.refEye = lightEye;
    root.refAt = lightAt;
    GI_TRY(evaluateAreaLightSource(root, root.level));
    GI_TRY(renderAreaLightSource(root, root.level));
    const int n = adaptiveSamples_;
    // configureShadow (MyGLGeometryViewer.cpp:184-194): column 3 of B * lightMVPs[i],
    // w += accFactor * 10000; columns 0-2 from B * the last light rendered
    std::vector<float> trans(static_cast<size_t>(n) * 4);
    float B[16], biased[16], common[16];
    GI_TRY(gi_bias_matrix(0.0f, B));
    for (int i = 0; i < n; ++i) {
        GI_TRY(gi_mat4_mul(B, lightMVPs_[i].data(), biased));
        for (int r = 0; r < 4; ++r) trans[4 * i + r] = biased[12 + r];
        trans[4 * i + 3] = trans[4 * i + 3] + accFactor_[i] * 10000.0f;
    }
    GI_TRY(gi_mat4_mul(B, lightMVP_.data(), common));
    GI_TRY(hip_ok(hipMemcpyAsync(trans_.get(), trans.data(), trans.size() * sizeof(float), hipMemcpyHostToDevice,
                                 stream)));
    if (revectorizationBasedAdaptiveSampling) {
        // RevectorizationBasedAccurateSoftShadow.frag (:868-871); the final
        // rendering turns the light 180 degrees about y (:365-368)
        gi_rev_soft_params rp{};
        rp.shadow_intensity = shadowIntensity;
        rp.rpcf = RPCF ? 1 : 0;
        rp.max_search = maxSearch;
        rp.depth_threshold = depthThreshold;
        std::memcpy(rp.mv, cameraMV_.data(), sizeof(rp.mv));
        normal_matrix(cameraMV_, rp.normal_matrix);
        rp.light_position[0] = -lightEye[0];
        rp.light_position[1] = lightEye[1];
        rp.light_position[2] = -lightEye[2];
        gi_gbuffer g{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
        GI_TRY(gi_rev_adaptive_soft_shadow_pass(&g, maps_.as<float>(), n, smw_, smh_, common, trans_.as<float>(),
                                                disc_.as<unsigned char>(), GI_REV_DISC_LAYERS, qtVis_.as<float>(),
                                                &rp, soft_.as<float>(), ws_.get(), ws_.size(), stream));
        // the next frame's RPCF (main.cpp:880-885)
        if (RPCF && n <= 9)
            RPCF = false;
        else if (!RPCF && n >= 81)
            RPCF = true;
    } else {
        gi_gbuffer g{gpos_.as<float>(), nullptr, windowWidth_, windowHeight_, nullptr};
        gi_soft_params sp{shadowIntensity, adaptiveSamplingLowerAccuracy ? 1 : 0, 3, 1.0f, 1};
        GI_TRY(gi_adaptive_soft_shadow_pass(&g, maps_.as<float>(), n, smw_, smh_, common, trans_.as<float>(), &sp,
                                            soft_.as<float>(), stream));
    }
    return hip_ok(hipStreamSynchronize(stream));  // the host trans buffer goes out of scope
}

std::array<float, 9> SoftShadowPipeline::normalMatrix() const {
    std::array<float, 9> nm{};
    normal_matrix(cameraMV_, nm.data());
    return nm;
}

int SoftShadowPipeline::deferredShading() {
    gi_gbuffer g{gpos_.as<float>(), gnrm_.as<float>(), windowWidth_, windowHeight_, gfac_.as<float>()};
    return phong(g, cameraMV_, lightEye, shadowIntensity, soft_.as<float>(), rgba_.as<float>(), stream);
}

int SoftShadowPipeline::download(std::vector<float>& softShadow, std::vector<float>* maps, std::vector<float>* rgba) {
    const size_t npx = static_cast<size_t>(windowWidth_) * windowHeight_;
    softShadow.resize(npx);
    GI_TRY(hip_ok(hipMemcpyAsync(softShadow.data(), soft_.get(), npx * sizeof(float), hipMemcpyDeviceToHost,
                                 stream)));
    if (maps) {
        maps->resize(static_cast<size_t>(numberOfSamples) * smw_ * smh_);
        GI_TRY(hip_ok(hipMemcpyAsync(maps->data(), maps_.get(), maps->size() * sizeof(float),
                                     hipMemcpyDeviceToHost, stream)));
    }
    if (rgba) {
        rgba->resize(4 * npx);
        GI_TRY(hip_ok(hipMemcpyAsync(rgba->data(), rgba_.get(), 4 * npx * sizeof(float), hipMemcpyDeviceToHost,
                                     stream)));
    }
    return hip_ok(hipStreamSynchronize(stream));
}

}  // namespace gi::host
