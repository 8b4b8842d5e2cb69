// host_pipeline_test.cpp — runs the C++ pass interface (shadow_passes.hpp)
// the way the reference apps' display() does, and dumps every input and
// output as raw little-endian files for tests/test_gpu_host_pipeline.py to
// check against the CPU oracle.
//
// usage: host_pipeline_test <out_dir>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../globalillumination_amd/host/shadow_passes.hpp"

using gi::host::HardShadowPipeline;
using gi::host::Mesh;
using gi::host::SoftShadowPipeline;

namespace {

// UV sphere (radius r, 64 x 32, centre c) + ground quad (20 x 20 at y = 0):
// the hard-shadow scene of the ShadowMapping app's sphere configuration.
Mesh scene() {
    Mesh m;
    const int seg = 64, rings = 32;
    const double pi = 3.14159265358979323846;
    auto add = [&](double x, double y, double z, double nx, double ny, double nz) {
        m.positions.insert(m.positions.end(), {float(x), float(y), float(z)});
        m.normals.insert(m.normals.end(), {float(nx), float(ny), float(nz)});
    };
    const double cy = 1.2;
    add(0, cy + 1, 0, 0, 1, 0);
    for (int r = 1; r < rings; ++r) {
        double th = pi * r / rings;
        for (int k = 0; k < seg; ++k) {
            double ph = 2 * pi * k / seg;
            double x = std::sin(th) * std::cos(ph), y = std::cos(th), z = std::sin(th) * std::sin(ph);
            add(x, cy + y, z, x, y, z);
        }
    }
    add(0, cy - 1, 0, 0, -1, 0);
    const int top = 0, bot = 1 + (rings - 1) * seg;
    auto tri = [&](int a, int b, int c) { m.indices.insert(m.indices.end(), {a, b, c}); };
    for (int k = 0; k < seg; ++k) tri(top, 1 + k, 1 + (k + 1) % seg);
    for (int r = 0; r < rings - 2; ++r)
        for (int k = 0; k < seg; ++k) {
            int a0 = 1 + r * seg, b0 = 1 + (r + 1) * seg, k1 = (k + 1) % seg;
            tri(a0 + k, b0 + k, a0 + k1);
            tri(a0 + k1, b0 + k, b0 + k1);
        }
    for (int k = 0; k < seg; ++k) tri(bot, 1 + (rings - 2) * seg + (k + 1) % seg, 1 + (rings - 2) * seg + k);
    int base = m.nverts();
    for (auto [x, z] : {std::pair{-10, -10}, {10, -10}, {10, 10}, {-10, 10}}) add(x, 0, z, 0, 1, 0);
    tri(base, base + 2, base + 1);
    tri(base, base + 3, base + 2);
    return m;
}

bool dump(const std::string& path, const void* data, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

template <typename T>
bool dumpv(const std::string& dir, const char* name, const std::vector<T>& v) {
    return dump(dir + "/" + name, v.data(), v.size() * sizeof(T));
}

int fail(const char* what, int rc) {
    std::fprintf(stderr, "%s failed: %s (%d)\n", what, gi_status_string(rc), rc);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const Mesh mesh = scene();
    if (!dumpv(dir, "mesh_pos.f32", mesh.positions) || !dumpv(dir, "mesh_nrm.f32", mesh.normals) ||
        !dumpv(dir, "mesh_idx.i32", mesh.indices))
        return fail("dump", -1);

    const int W = 320, H = 240, SM = 512;
    HardShadowPipeline hard;
    hard.lightEye = {4.0f, 8.0f, 4.0f};
    hard.lightAt = {0.0f, 0.0f, 0.0f};
    hard.cameraEye = {0.0f, 3.0f, 6.0f};
    hard.cameraAt = {0.0f, 0.0f, 0.0f};
    hard.cameraFov = 60.0f;
    hard.shadowParams.zBias = 0.005f;
    hard.shadowParams.kernelOrder = 3;
    hard.shadowParams.depthThreshold = 2.5e-6f;
    int rc = hard.init(mesh, W, H, SM, SM);
    if (rc) return fail("HardShadowPipeline::init", rc);

    struct Technique {
        const char* name;
        bool naive, pcf, smsr, rpcf, rsmss, conservative;
    };
    const Technique techs[] = {{"naive", true, false, false, false, false, false},
                               {"pcf", false, true, false, false, false, false},
                               {"rbsm", false, false, true, false, false, false},
                               {"rpcf", false, false, false, true, false, false},
                               {"rbsm_filtered", false, false, false, false, true, false},
                               {"rbsm_conservative", false, false, true, false, false, true},
                               {"rpcf_conservative", false, false, false, true, false, true}};
    bool first = true;
    for (const Technique& t : techs) {
        auto& p = hard.shadowParams;
        p.naive = t.naive;
        p.bilinearPCF = t.pcf;
        p.SMSR = t.smsr;
        p.RPCFPlusSMSR = t.rpcf;
        p.RSMSS = t.rsmss;
        p.conservative = t.conservative;
        rc = hard.display();
        if (rc) return fail(t.name, rc);
        std::vector<float> out, depth, gpos, gnrm, rgba;
        rc = first ? hard.download(out, &depth, &gpos, &gnrm, &rgba) : hard.download(out, nullptr, nullptr, nullptr, &rgba);
        if (rc) return fail("download", rc);
        if (first) {
            dumpv(dir, "depth.f32", depth);
            dumpv(dir, "gbuf_pos.f32", gpos);
            dumpv(dir, "gbuf_nrm.f32", gnrm);
            dump(dir + "/light_mvp.f32", p.lightMVP.data(), 64);
            dump(dir + "/camera_mvp.f32", hard.cameraMVP().data(), 64);
            dump(dir + "/camera_mv.f32", hard.cameraMV().data(), 64);
            dump(dir + "/camera_nm.f32", hard.normalMatrix().data(), 36);
            first = false;
        }
        dumpv(dir, (std::string("hard_") + t.name + ".f32").c_str(), out);
        dumpv(dir, (std::string("phong_") + t.name + ".f32").c_str(), rgba);
    }

    SoftShadowPipeline soft;
    soft.lightEye = {4.0f, 8.0f, 4.0f};
    soft.lightAt = {0.0f, 0.0f, 0.0f};
    soft.lightSize = 1;
    soft.numberOfSamples = 9;
    soft.adaptiveSamplingLowerAccuracy = true;
    soft.cameraEye = {0.0f, 3.0f, 6.0f};
    soft.cameraAt = {0.0f, 0.0f, 0.0f};
    soft.cameraFov = 60.0f;
    rc = soft.init(mesh, W, H, SM, SM, 289);  // 289: adaptive sampling's 17 x 17 grid
    if (rc) return fail("SoftShadowPipeline::init", rc);
    rc = soft.renderGBuffer();
    if (rc) return fail("SoftShadowPipeline::renderGBuffer", rc);
    rc = soft.renderMonteCarlo();
    if (rc) return fail("renderMonteCarlo", rc);
    rc = soft.deferredShading();
    if (rc) return fail("deferredShading", rc);
    std::vector<float> sout, maps, srgba;
    rc = soft.download(sout, &maps, &srgba);
    if (rc) return fail("download", rc);
    dumpv(dir, "soft.f32", sout);
    dumpv(dir, "soft_phong.f32", srgba);
    dump(dir + "/soft_camera_mv.f32", soft.cameraMV().data(), 64);
    dump(dir + "/soft_camera_nm.f32", soft.normalMatrix().data(), 36);
    dumpv(dir, "soft_maps.f32", maps);
    dump(dir + "/soft_light_mvps.f32", soft.lightMVPs().data(), soft.lightMVPs().size() * 64);

    // renderSoftShadows (PCSS) from a far light: window depths above the
    // shader's 0.99 blocker cut-off
    soft.lightEye = {60.0f, 120.0f, 60.0f};
    soft.lightSize = 16;
    rc = soft.renderSoftShadows();
    if (rc) return fail("renderSoftShadows", rc);
    std::vector<float> pcss;
    rc = soft.download(pcss, &maps);
    if (rc) return fail("download", rc);
    maps.resize(static_cast<size_t>(SM) * SM);
    dumpv(dir, "pcss.f32", pcss);
    dumpv(dir, "pcss_map.f32", maps);
    dump(dir + "/pcss_light_mvp.f32", soft.lightMVP().data(), 64);

    // the same frame with ShadowParams.RBSSM: kernel 7, maxSearch 16
    soft.rbssm = true;
    soft.kernelSize = 7;
    soft.depthThreshold = 2.5e-5f;
    rc = soft.renderSoftShadows();
    if (rc) return fail("renderSoftShadows(RBSSM)", rc);
    std::vector<float> rbssm;
    rc = soft.download(rbssm);
    if (rc) return fail("download", rc);
    dumpv(dir, "rbssm.f32", rbssm);
    soft.rbssm = false;
    soft.kernelSize = 15;

    // renderAdaptiveLightSourceSampling: a 2 x 2 light at (4, 8, 4), maximum
    // level 4, exact evaluation (div 1) and the lower-accuracy variant
    // (4x4 PCF per sample, evaluation on a quarter-resolution grid)
    soft.lightEye = {4.0f, 8.0f, 4.0f};
    soft.lightSize = 2;
    for (int lower = 0; lower < 2; ++lower) {
        soft.adaptiveSamplingLowerAccuracy = lower != 0;
        rc = soft.renderAdaptiveLightSourceSampling();
        if (rc) return fail("renderAdaptiveLightSourceSampling", rc);
        std::vector<float> ad;
        rc = soft.download(ad);
        if (rc) return fail("download", rc);
        const std::string tag = lower ? "adaptive_lower" : "adaptive";
        dumpv(dir, (tag + "_soft.f32").c_str(), ad);
        dumpv(dir, (tag + "_acc.f32").c_str(), soft.accFactors());
        dump(dir + "/" + tag + "_mvps.f32", soft.lightMVPs().data(), soft.lightMVPs().size() * 64);
        std::vector<int32_t> tests;
        for (const auto& t : soft.nodeTests())
            tests.insert(tests.end(), {t.nodeIndex, t.level, static_cast<int32_t>(t.count)});
        dumpv(dir, (tag + "_tests.i32").c_str(), tests);
        std::printf("%s: %d samples, %zu node tests\n", tag.c_str(), soft.adaptiveSamples(), soft.nodeTests().size());
    }
    // the revectorization-based variant (menu entry 2): three frames over the
    // same discontinuity layers, the third with RPCF forced on
    soft.revectorizationBasedAdaptiveSampling = true;
    soft.adaptiveSamplingLowerAccuracy = true;
    soft.maxSearch = 16;
    soft.depthThreshold = 2.5e-6f;
    for (int frame = 0; frame < 3; ++frame) {
        if (frame == 2) soft.RPCF = true;
        const int32_t rpcf = soft.RPCF ? 1 : 0;
        rc = soft.renderAdaptiveLightSourceSampling();
        if (rc) return fail("renderAdaptiveLightSourceSampling(revectorization)", rc);
        std::vector<float> ad;
        rc = soft.download(ad);
        if (rc) return fail("download", rc);
        const std::string tag = "rev" + std::to_string(frame);
        dumpv(dir, (tag + "_soft.f32").c_str(), ad);
        dumpv(dir, (tag + "_acc.f32").c_str(), soft.accFactors());
        dump(dir + "/" + tag + "_mvps.f32", soft.lightMVPs().data(), soft.lightMVPs().size() * 64);
        std::vector<int32_t> tests{rpcf};
        for (const auto& t : soft.nodeTests())
            tests.insert(tests.end(), {t.nodeIndex, t.level, static_cast<int32_t>(t.count)});
        dumpv(dir, (tag + "_tests.i32").c_str(), tests);
        std::printf("%s: %d samples, %zu node tests, RPCF %d\n", tag.c_str(), soft.adaptiveSamples(),
                    soft.nodeTests().size(), rpcf);
    }
    std::printf("host pipeline ok: %d vertices, %d triangles\n", mesh.nverts(), mesh.ntris());
    return 0;
}
