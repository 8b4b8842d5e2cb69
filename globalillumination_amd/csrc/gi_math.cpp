// gi_math.cpp — host-side light camera math of the C ABI (include/gi.h).
//
// FP32 restatement of the three glm 0.9.3.1 formulas the reference uses to
// build light and camera matrices (MyGLGeometryViewer::configureAmbient,
// ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:14-26) and of the
// texture-space bias of configureShadow (:139-147).  Built without FMA
// contraction so the FP32 results are those of the formulas as written.
#include <cmath>
#include <cstring>

#include "../../include/gi.h"

namespace {

bool finite_all(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

struct V3 { float x, y, z; };

V3 normalize(V3 v) {
    // glm/core/func_geometric.inl:239-248, inversesqrt(x) = 1/sqrt(x)
    float sqr = v.x * v.x + v.y * v.y + v.z * v.z;
    float inv = 1.0f / std::sqrt(sqr);
    return {v.x * inv, v.y * inv, v.z * inv};
}

V3 cross(V3 a, V3 b) {
    // glm/core/func_geometric.inl:199-211
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

}  // namespace

extern "C" {

int gi_api_version(void) { return GI_API_VERSION; }

const char* gi_status_string(int status) {
    switch (status) {
    case GI_OK: return "ok";
    case GI_EINVAL: return "invalid argument";
    case GI_ENONFINITE: return "non-finite parameter";
    case GI_EDEVICE: return "HIP device error";
    case GI_ENOSPACE: return "workspace too small";
    case GI_ENOTSUP: return "size not supported";
    case GI_EIO: return "file cannot be opened or parsed";
    default: return "unknown status";
    }
}

int gi_perspective(float fovy, float aspect, float zn, float zf, float out[16]) {
    if (!out) return GI_EINVAL;
    if (!std::isfinite(fovy) || !std::isfinite(aspect) || !std::isfinite(zn) || !std::isfinite(zf))
        return GI_ENONFINITE;
    if (!(aspect > 0.0f) || !(zn > 0.0f) || !(zf > zn) || !(fovy > 0.0f && fovy < 180.0f))
        return GI_EINVAL;
    // glm/gtc/matrix_transform.inl:223-244, radians() = deg * (pi / 180)
    const float pi = 3.1415926535897932384626433832795f;
    float range = std::tan((fovy / 2.0f) * (pi / 180.0f)) * zn;
    float left = -range * aspect, right = range * aspect;
    float bottom = -range, top = range;
    std::memset(out, 0, 16 * sizeof(float));
    out[0] = (2.0f * zn) / (right - left);
    out[5] = (2.0f * zn) / (top - bottom);
    out[10] = -(zf + zn) / (zf - zn);
    out[11] = -1.0f;
    out[14] = -(2.0f * zf * zn) / (zf - zn);
    return GI_OK;
}

int gi_ortho(float l, float r, float b, float t, float zn, float zf, float out[16]) {
    if (!out) return GI_EINVAL;
    float in[6] = {l, r, b, t, zn, zf};
    if (!finite_all(in, 6)) return GI_ENONFINITE;
    if (r == l || t == b || zf == zn) return GI_EINVAL;
    // glm/gtc/matrix_transform.inl:163-182
    std::memset(out, 0, 16 * sizeof(float));
    out[0] = 2.0f / (r - l);
    out[5] = 2.0f / (t - b);
    out[10] = -2.0f / (zf - zn);
    out[12] = -(r + l) / (r - l);
    out[13] = -(t + b) / (t - b);
    out[14] = -(zf + zn) / (zf - zn);
    out[15] = 1.0f;
    return GI_OK;
}

int gi_lookat(const float eye[3], const float center[3], const float up[3], float out[16]) {
    if (!eye || !center || !up || !out) return GI_EINVAL;
    if (!finite_all(eye, 3) || !finite_all(center, 3) || !finite_all(up, 3)) return GI_ENONFINITE;
    // glm/gtc/matrix_transform.inl:383-411
    V3 f = normalize({center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]});
    V3 u = normalize({up[0], up[1], up[2]});
    V3 s = normalize(cross(f, u));
    u = cross(s, f);
    float rot[16] = {s.x, u.x, -f.x, 0.0f, s.y, u.y, -f.y, 0.0f,
                     s.z, u.z, -f.z, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
    if (!finite_all(rot, 16)) return GI_EINVAL;  // eye == center or up parallel to the view axis
    // translate(rot, -eye) (:32-41): column 3 = r0*(-ex) + r1*(-ey) + r2*(-ez) + r3
    float tx = -eye[0], ty = -eye[1], tz = -eye[2];
    std::memcpy(out, rot, sizeof(rot));
    for (int i = 0; i < 4; ++i) {
        float acc = rot[i] * tx;
        acc = acc + rot[4 + i] * ty;
        acc = acc + rot[8 + i] * tz;
        acc = acc + rot[12 + i];
        out[12 + i] = acc;
    }
    return GI_OK;
}

int gi_mat4_mul(const float a[16], const float b[16], float out[16]) {
    if (!a || !b || !out) return GI_EINVAL;
    // glm/core/type_mat4x4.inl:757-777
    float r[16];
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i) {
            float acc = a[i] * b[c * 4 + 0];
            acc = acc + a[4 + i] * b[c * 4 + 1];
            acc = acc + a[8 + i] * b[c * 4 + 2];
            acc = acc + a[12 + i] * b[c * 4 + 3];
            r[c * 4 + i] = acc;
        }
    std::memcpy(out, r, sizeof(r));
    return GI_OK;
}

int gi_bias_matrix(float zbias, float out[16]) {
    if (!out) return GI_EINVAL;
    if (!std::isfinite(zbias)) return GI_ENONFINITE;
    // ShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:139-145
    std::memset(out, 0, 16 * sizeof(float));
    out[0] = 0.5f;
    out[5] = 0.5f;
    out[10] = 0.5f;
    out[12] = 0.5f;
    out[13] = 0.5f;
    out[14] = 0.5f - zbias;
    out[15] = 1.0f;
    return GI_OK;
}

int gi_uniform_sampling(const float base[3], int size, int nsamples, int index, float out[3]) {
    if (!base || !out || nsamples <= 1 || index < 0 || index >= nsamples) return GI_EINVAL;
    // SoftShadowMapping/src/Scene/LightSource/UniformSampledLightSource.cpp:27-37;
    // halfSize and sampleSize go through a double divide as in the source.
    float halfSize = static_cast<float>(static_cast<double>(static_cast<float>(size)) / 2.0);
    float factor = std::sqrt(static_cast<float>(nsamples));
    float sampleSize = static_cast<float>(static_cast<double>(factor - 1.0f) / 2.0);
    if (!(sampleSize > 0.0f)) return GI_EINVAL;
    out[0] = base[0];
    out[1] = base[1];
    out[2] = base[2];
    out[0] += ((static_cast<float>(index % static_cast<int>(factor)) - sampleSize) / sampleSize) * halfSize;
    out[1] += ((static_cast<float>(static_cast<int>(static_cast<float>(index) / factor)) - sampleSize) /
               sampleSize) * halfSize;
    return GI_OK;
}

int gi_soft_light_trans(const float* light_mvps, int nsamples, float* trans_out) {
    if (!light_mvps || !trans_out || nsamples <= 0 || nsamples > GI_MAX_SAMPLES) return GI_EINVAL;
    // SoftShadowMapping/src/Viewers/MyGLGeometryViewer.cpp:184-195
    for (int s = 0; s < nsamples; ++s)
        for (int r = 0; r < 4; ++r) trans_out[4 * s + r] = light_mvps[16 * s + 12 + r];
    return GI_OK;
}

}  // extern "C"
