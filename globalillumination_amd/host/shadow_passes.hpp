// shadow_passes.hpp — C++ host pass interface over the C ABI of include/gi.h.
//
// Mirrors the structure of the reference apps: the ShadowParams struct
// (ShadowMapping/include/Viewers/ShadowParams.h:6-37), the light/camera setup
// of MyGLGeometryViewer (ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:6-26)
// and the pass functions of ShadowMapping/src/main.cpp (renderShadowMap :350,
// renderGBuffer :363, computeHardShadows :400, display :459) and
// SoftShadowMapping/src/main.cpp (renderMonteCarlo :756).  Where the reference
// binds GL textures, these objects own device buffers allocated once in init()
// (as initGL allocates its textures, main.cpp:886-899); every pass is then a
// libgi kernel launch on the pipeline's stream.  Passes return a gi_status.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <vector>

#include "../../include/gi.h"

namespace gi::host {

struct Mesh {
    std::vector<float> positions;  // nverts x 3
    std::vector<float> normals;    // nverts x 3
    std::vector<int32_t> indices;  // ntris x 3
    int nverts() const { return static_cast<int>(positions.size() / 3); }
    int ntris() const { return static_cast<int>(indices.size() / 3); }
};

using Mat4 = std::array<float, 16>;  // column-major, glm layout
using Vec3 = std::array<float, 3>;

// ShadowMapping/include/Viewers/ShadowParams.h:6-37 (GL texture handles are
// the pipeline's device buffers).
struct ShadowParams {
    Mat4 lightMVP{};
    Mat4 lightMV{};
    Mat4 lightP{};
    int shadowMapWidth = 0;
    int shadowMapHeight = 0;
    int maxSearch = 16;       // main.cpp:872
    int kernelOrder = 7;      // gaussianFilter->getOrder(), main.cpp:859
    int penumbraSize = 1;     // main.cpp:871
    float depthThreshold = 0.0f;
    float shadowIntensity = 0.25f;  // main.cpp:877
    bool tricubicPCF = false;
    bool bilinearPCF = false;
    bool naive = true;        // main.cpp:876 default
    bool SMSR = false;
    bool RPCFPlusSMSR = false;
    bool RSMSS = false;
    bool RPCFPlusRSMSS = false;
    bool conservative = false;
    bool preEvaluation = true;  // computePreEvaluationBasedOnNormalOrientation
    float zBias = 0.0f;         // constant bias folded into B[3][2]
};

// MyGLGeometryViewer: fov 45, near 1, far 1000 (:3-12); configureAmbient
// builds P = perspective(fov, W/H, near, far), V = lookAt(eye, look, up), M = I.
struct GeometryViewer {
    float fov = 45.0f, zNear = 1.0f, zFar = 1000.0f;
    Vec3 eye{}, look{}, up{0.0f, 0.0f, 1.0f};
    Mat4 projection{}, view{}, model{};
    int configureAmbient(int width, int height);
    Mat4 mvp() const;  // projection * view * model
};

// G-buffer colour of a mesh without per-vertex colours: GBuffer.frag:27 reads
// gl_FrontLightModelProduct.sceneColor = emission (0) + the default material
// ambient (0.2) times the light-model ambient (0.1, MyGLGeometryViewer.cpp:69),
// alpha 1 (the default material diffuse alpha).
constexpr float kSceneColor[4] = {0.2f * 0.1f, 0.2f * 0.1f, 0.2f * 0.1f, 1.0f};
// shadeScene / deferredShading clear colour (main.cpp:452).
constexpr float kClearColor[4] = {0.63f, 0.82f, 0.96f, 1.0f};

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    ~DeviceBuffer();
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    int allocate(size_t bytes);
    void* get() const { return ptr_; }
    size_t size() const { return bytes_; }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr_); }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// The hard-shadow frame of ShadowMapping's display() (main.cpp:459-472):
// renderShadowMap -> renderGBuffer -> computeHardShadows.
class HardShadowPipeline {
public:
    ShadowParams shadowParams;
    Vec3 lightEye{}, lightAt{}, lightUp{0.0f, 0.0f, 1.0f};     // main.cpp:866-867
    Vec3 cameraEye{}, cameraAt{}, cameraUp{0.0f, 0.0f, 1.0f};
    float cameraFov = 45.0f;
    gi_raster_opts rasterOpts{4.0f, 20.0f, nullptr, nullptr};  // glPolygonOffset(4, 20), main.cpp:246
    hipStream_t stream = nullptr;

    ~HardShadowPipeline();
    // initGL (main.cpp:839-953): device buffers for the map, G-buffer, output.
    int init(const Mesh& mesh, int windowWidth, int windowHeight, int shadowMapWidth, int shadowMapHeight);
    int renderShadowMap();     // main.cpp:350-361
    int renderGBuffer();       // main.cpp:363-372
    int computeHardShadows();  // main.cpp:400-414
    int shadeScene();          // main.cpp:449-456: deferred Phong over the hard-shadow buffer
    int display();             // main.cpp:459-472 without the GL swap
    int download(std::vector<float>& hardShadow, std::vector<float>* depth = nullptr,
                 std::vector<float>* gbufferPosition = nullptr, std::vector<float>* gbufferNormal = nullptr,
                 std::vector<float>* rgba = nullptr);
    const Mat4& cameraMVP() const { return cameraMVP_; }
    const Mat4& cameraMV() const { return cameraMV_; }
    std::array<float, 9> normalMatrix() const;  // inverseTranspose(mat3(MV)), configurePhong

private:
    int technique() const;
    int windowWidth_ = 0, windowHeight_ = 0, nverts_ = 0, ntris_ = 0;
    bool ownStream_ = false;
    Mat4 cameraMVP_{}, cameraMV_{};
    DeviceBuffer pos_, nrm_, idx_, mvp_, depth_, gpos_, gnrm_, gfac_, hard_, rgba_, ws_;
};

// The Monte-Carlo area light of SoftShadowMapping (renderMonteCarlo,
// SoftShadowMapping/src/main.cpp:756-811): one depth map per point-light
// sample, then AccurateSoftShadow over the G-buffer.
class SoftShadowPipeline {
public:
    float shadowIntensity = 0.25f;                 // Soft main.cpp:1611
    bool adaptiveSamplingLowerAccuracy = false;    // 4x4 PCF per sample when set
    Vec3 lightEye{}, lightAt{}, lightUp{0.0f, 0.0f, 1.0f};
    int lightSize = 16;                            // LightSource::size (int), main.cpp:1598
    int numberOfSamples = 289;                     // main.cpp:1600
    int blockerSearchSize = 7;                     // main.cpp:1612
    int kernelSize = 15;                           // main.cpp:1613
    // The int uniform zNear of PlausibleSoftShadow.frag.  The reference never
    // uploads it for that shader, so it is 0 there and PCSS degenerates to
    // "lit" (penumbra width 0); 1 is MyGLGeometryViewer's near plane.
    int pcssZNear = 1;
    // ShadowParams.RBSSM (menu entry 5, main.cpp:1451-1455, kernelSize 7):
    // renderSoftShadows runs RBSSM.frag instead of PlausibleSoftShadow.frag
    // (main.cpp:1012-1013) with the edge walk of maxSearch texels.
    bool rbssm = false;
    int maxSearch = 16;                            // main.cpp:1617
    float depthThreshold = 0.0f;                   // SceneLoader `d`, main.cpp:1618
    int maxLevel = 4;                              // main.cpp:190
    // Menu entry 2 (main.cpp:1411-1416): discontinuity layers per evaluated
    // corner map, the revectorized node test and final pass.  RPCF is the
    // frame-to-frame state the reference keeps in ShadowParams.RPCF: after
    // each frame it turns on at 81 or more samples and off at 9 or fewer.
    bool revectorizationBasedAdaptiveSampling = false;
    bool RPCF = false;
    bool preEvaluation = true;                     // PCSS runs the pre-evaluation (:614)
    Vec3 cameraEye{}, cameraAt{}, cameraUp{0.0f, 0.0f, 1.0f};
    float cameraFov = 45.0f;
    hipStream_t stream = nullptr;

    ~SoftShadowPipeline();
    int init(const Mesh& mesh, int windowWidth, int windowHeight, int shadowMapWidth, int shadowMapHeight,
             int maxSamples);
    int renderGBuffer();
    // UniformSampledLightSource placement (eye and look-at shifted together,
    // UniformSampledLightSource.cpp:13-37), the N depth maps, then the screen pass.
    int renderMonteCarlo();
    // renderSoftShadows with ShadowParams.PCSS or .RBSSM (main.cpp:925-1016):
    // the depth map of the light centre (into map 0), then
    // PlausibleSoftShadow.frag or RBSSM.frag.
    int renderSoftShadows();
    // renderAdaptiveLightSourceSampling (main.cpp:813-892): the quad tree
    // over the light is refined where a node's four corner maps disagree on
    // a penumbra pixel (evaluateAreaLightSource :573-620), every leaf's
    // corner maps are rendered (renderAreaLightSource :622-694), and the
    // weighted AccurateSoftShadow pass averages them (or, with
    // revectorizationBasedAdaptiveSampling, the revectorized passes).
    // Needs maxSamples >= 289.
    int renderAdaptiveLightSourceSampling();
    // The adaptive frame's record: node tests in evaluation order
    // (node index, level, passing pixels) and each map's quad-tree level.
    struct NodeTest {
        int nodeIndex, level;
        unsigned count;
    };
    const std::vector<NodeTest>& nodeTests() const { return nodeTests_; }
    const std::vector<float>& accFactors() const { return accFactor_; }
    int adaptiveSamples() const { return adaptiveSamples_; }
    int deferredShading();  // Soft main.cpp:1109-1117: Phong over the soft-shadow buffer
    int download(std::vector<float>& softShadow, std::vector<float>* maps = nullptr,
                 std::vector<float>* rgba = nullptr);
    const std::vector<Mat4>& lightMVPs() const { return lightMVPs_; }
    const Mat4& cameraMVP() const { return cameraMVP_; }
    const Mat4& cameraMV() const { return cameraMV_; }
    std::array<float, 9> normalMatrix() const;
    const Mat4& lightMVP() const { return lightMVP_; }  // the PCSS light centre's P*V

private:
    int windowWidth_ = 0, windowHeight_ = 0, smw_ = 0, smh_ = 0, maxSamples_ = 0, nverts_ = 0, ntris_ = 0;
    bool ownStream_ = false;
    std::vector<Mat4> lightMVPs_;
    Mat4 cameraMVP_{}, cameraMV_{}, lightMVP_{};
    DeviceBuffer pos_, nrm_, idx_, mvps_, trans_, maps_, gpos_, gnrm_, gfac_, soft_, rgba_, ws_, qtImg_, qtCount_;
    DeviceBuffer qtVis_, disc_;  // revectorization-based adaptive sampling
    // adaptive sampling state (renderAdaptiveLightSourceSampling)
    struct QuadNode;
    int renderLightSample(const Vec3& eye, const Vec3& at);
    int evaluateSubAreaLightSource(const QuadNode& node, int nodeIndex, bool& subdivide);
    int evaluateAreaLightSource(QuadNode& node, int nodeIndex);
    int renderAreaLightSource(const QuadNode& node, int nodeIndex);
    bool qtRendered_[17][17]{};
    int qtHash_[17][17]{};
    std::vector<NodeTest> nodeTests_;
    std::vector<float> accFactor_;
    int adaptiveSamples_ = 0;
};

}  // namespace gi::host
