// scene_loader.cpp — SceneLoader / Mesh / OBJ reading for the shadow pipeline,
// and the C ABI of include/gi_scene.h.  Plain host C++; built with
// -ffp-contract=off so each FP32 expression rounds where the reference's does.
#include "scene_loader.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>

#include "../../include/gi_scene.h"

namespace gi::host {
namespace {

// One face-vertex reference as glmSecondPass reads it (OBJLoader.cpp:651-760):
// the format of the face's first token ("v", "v/t", "v//n", "v/t/n") fixes how
// the others are read; negative indices count back from the current end.
enum class FaceFormat { V, VT, VN, VTN };

FaceFormat face_format(const std::string& tok) {
    if (tok.find("//") != std::string::npos) return FaceFormat::VN;
    int slashes = 0;
    for (char c : tok) slashes += (c == '/');
    if (slashes == 2) return FaceFormat::VTN;
    if (slashes == 1) return FaceFormat::VT;
    return FaceFormat::V;
}

// Parses one face token in the face's format; false if it does not match
// (glm's `while (fscanf(...) > 0)` stops there).
bool face_ref(const std::string& tok, FaceFormat f, long& v, long& t) {
    const char* s = tok.c_str();
    char* end = nullptr;
    v = std::strtol(s, &end, 10);
    if (end == s) return false;
    t = 0;
    if (f == FaceFormat::V) return *end == '\0';
    if (*end != '/') return false;
    if (f == FaceFormat::VN) return end[1] == '/';  // the normal index is not used
    const char* ts = end + 1;
    t = std::strtol(ts, &end, 10);
    return end != ts;
}

// glm's 1-based index with negative values relative to the current count
// (`v < 0 ? v + numvertices : v`, numvertices starting at 1).
long resolve(long idx, long count) { return idx < 0 ? idx + count + 1 : idx; }

float parse_float(const std::string& s, bool& ok) {
    const char* p = s.c_str();
    char* end = nullptr;
    float v = std::strtof(p, &end);
    ok = end != p;
    return v;
}

}  // namespace

int SceneMesh::loadOBJFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) return -1;
    std::vector<float> verts, norms, tex;  // glm keeps 1-based arrays; these are 0-based
    struct Tri {
        long v[3], t[3];
    };
    std::vector<Tri> tris;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "v" || key == "vn") {
            float xyz[3] = {0.0f, 0.0f, 0.0f};
            for (float& c : xyz) {
                std::string tok;
                bool ok = false;
                if (ls >> tok) c = parse_float(tok, ok);
                if (!ok) return -1;
            }
            auto& dst = (key == "v") ? verts : norms;
            dst.insert(dst.end(), xyz, xyz + 3);
        } else if (key == "vt") {
            float uv[2] = {0.0f, 0.0f};
            for (float& c : uv) {
                std::string tok;
                bool ok = false;
                if (ls >> tok) c = parse_float(tok, ok);
                if (!ok) return -1;
            }
            tex.insert(tex.end(), uv, uv + 2);
        } else if (key == "f") {
            std::vector<std::string> toks;
            for (std::string tok; ls >> tok;) toks.push_back(tok);
            if (toks.size() < 3) return -1;
            const FaceFormat fmt = face_format(toks[0]);
            const long nv = static_cast<long>(verts.size() / 3), nt = static_cast<long>(tex.size() / 2);
            long v[3], t[3];
            for (int k = 0; k < 3; ++k) {
                if (!face_ref(toks[k], fmt, v[k], t[k])) return -1;
                v[k] = resolve(v[k], nv);
                t[k] = resolve(t[k], nt);
            }
            tris.push_back({{v[0], v[1], v[2]}, {t[0], t[1], t[2]}});
            // fan: (first, previous last, new), OBJLoader.cpp:667-676
            for (size_t k = 3; k < toks.size(); ++k) {
                long vn, tn;
                if (!face_ref(toks[k], fmt, vn, tn)) break;
                const Tri& prev = tris.back();
                tris.push_back({{prev.v[0], prev.v[2], resolve(vn, nv)}, {prev.t[0], prev.t[2], resolve(tn, nt)}});
            }
        }
        // '#', g, usemtl, mtllib, s, o and anything else: the rest of the line is skipped
    }
    const long nverts = static_cast<long>(verts.size() / 3);
    for (const Tri& t : tris)
        for (long v : t.v)
            if (v < 1 || v > nverts) return -1;  // an index glm would read out of bounds
    positions = verts;
    // Mesh.cpp:274-283 copies normal i into vertex i when the file has normals;
    // computeNormals then overwrites every vertex a triangle references.
    normals.assign(positions.size(), 0.0f);
    for (size_t i = 0; i < normals.size() && i < norms.size(); ++i) normals[i] = norms[i];
    indices.resize(tris.size() * 3);
    for (size_t i = 0; i < tris.size(); ++i)
        for (int k = 0; k < 3; ++k) indices[i * 3 + k] = static_cast<int32_t>(tris[i].v[k] - 1);
    textureCoords.assign(positions.size(), 0.0f);
    if (!tex.empty()) {  // Mesh.cpp:287-301: the last face corner to name a vertex wins
        const long ntex = static_cast<long>(tex.size() / 2);
        for (const Tri& t : tris)
            for (int k = 0; k < 3; ++k) {
                const long vc = t.v[k], tc = t.t[k];
                if (tc < 1 || tc > ntex) continue;
                textureCoords[(vc - 1) * 3 + 0] = tex[(tc - 1) * 2 + 0];
                textureCoords[(vc - 1) * 3 + 1] = tex[(tc - 1) * 2 + 1];
                textureCoords[(vc - 1) * 3 + 2] = 0.0f;
            }
    }
    colors.clear();
    textures.clear();
    textureIds = 0;
    return 0;
}

void SceneMesh::computeNormals() {
    const int n = nverts();
    normals.resize(positions.size(), 0.0f);
    std::vector<int> seen(n, 0);
    const float* P = positions.data();
    for (size_t i = 0; i + 2 < indices.size(); i += 3) {
        const int v[3] = {indices[i], indices[i + 1], indices[i + 2]};
        const float* a = P + 3 * v[0];
        const float* b = P + 3 * v[1];
        const float* c = P + 3 * v[2];
        const float x[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const float y[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        // glm::cross (func_geometric.inl:199-211) then glm::normalize: v * (1 / sqrt(dot(v, v)))
        const float cr[3] = {x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0], x[0] * y[1] - y[0] * x[1]};
        const float dot = (cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2];
        const float inv = 1.0f / std::sqrt(dot);
        const float nrm[3] = {cr[0] * inv, cr[1] * inv, cr[2] * inv};
        for (int j = 0; j < 3; ++j) {
            const int cur = v[j];
            const int k = ++seen[cur];
            for (int q = 0; q < 3; ++q) {
                float& dst = normals[3 * cur + q];
                if (k == 1) {
                    dst = nrm[q];
                } else {
                    // running mean in double: n * (1 - 1/k) + normal * 1 / k, stored as float
                    const double kd = static_cast<double>(k);
                    dst = static_cast<float>(static_cast<double>(dst) * (1.0 - 1.0 / kd) +
                                             static_cast<double>(nrm[q]) * 1.0 / kd);
                }
            }
        }
    }
}

void SceneMesh::scale(float x, float y, float z) {
    for (size_t i = 0; i + 2 < positions.size(); i += 3) {
        positions[i] *= x;
        positions[i + 1] *= y;
        positions[i + 2] *= z;
    }
}

void SceneMesh::translate(float x, float y, float z) {
    for (size_t i = 0; i + 2 < positions.size(); i += 3) {
        positions[i] += x;
        positions[i + 1] += y;
        positions[i + 2] += z;
    }
}

namespace {

using M4 = std::array<float, 16>;  // column-major, m[col * 4 + row]

// glm::rotate(mat4(1), angle, axis) with angles in degrees
// (gtc/matrix_transform.inl:44-78, radians() = deg * (pi / 180)).
M4 rotation(float deg, float ax, float ay, float az) {
    const float pi = 3.1415926535897932384626433832795f;
    const float a = deg * (pi / 180.0f);
    const float c = std::cos(a), s = std::sin(a);
    const float len2 = (ax * ax + ay * ay) + az * az;
    const float inv = 1.0f / std::sqrt(len2);
    const float axis[3] = {ax * inv, ay * inv, az * inv};
    const float temp[3] = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    float R[3][3];
    R[0][0] = c + temp[0] * axis[0];
    R[0][1] = (0.0f + temp[0] * axis[1]) + s * axis[2];
    R[0][2] = (0.0f + temp[0] * axis[2]) - s * axis[1];
    R[1][0] = (0.0f + temp[1] * axis[0]) - s * axis[2];
    R[1][1] = c + temp[1] * axis[1];
    R[1][2] = (0.0f + temp[1] * axis[2]) + s * axis[0];
    R[2][0] = (0.0f + temp[2] * axis[0]) + s * axis[1];
    R[2][1] = (0.0f + temp[2] * axis[1]) - s * axis[0];
    R[2][2] = c + temp[2] * axis[2];
    // Result[i] = m[0]*R[i][0] + m[1]*R[i][1] + m[2]*R[i][2] with m = identity
    M4 out{};
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 4; ++row) {
            const float m0 = (row == 0) ? 1.0f : 0.0f, m1 = (row == 1) ? 1.0f : 0.0f,
                        m2 = (row == 2) ? 1.0f : 0.0f;
            out[col * 4 + row] = (m0 * R[col][0] + m1 * R[col][1]) + m2 * R[col][2];
        }
    out[15] = 1.0f;
    return out;
}

// glm mat4 * mat4 (type_mat4x4.inl:757-777), columns summed left to right
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int c = 0; c < 4; ++c)
        for (int i = 0; i < 4; ++i) {
            float acc = a[i] * b[c * 4 + 0];
            acc = acc + a[4 + i] * b[c * 4 + 1];
            acc = acc + a[8 + i] * b[c * 4 + 2];
            acc = acc + a[12 + i] * b[c * 4 + 3];
            r[c * 4 + i] = acc;
        }
    return r;
}

}  // namespace

void SceneMesh::rotate(float x, float y, float z) {
    M4 R = mul(mul(rotation(x, 1.0f, 0.0f, 0.0f), rotation(y, 0.0f, 1.0f, 0.0f)), rotation(z, 0.0f, 0.0f, 1.0f));
    // transpose(R)[r][k] = R[k][r]: rotX = p.x*R[0][0] + p.y*R[1][0] + p.z*R[2][0]
    auto apply = [&](std::vector<float>& v) {
        for (size_t i = 0; i + 2 < v.size(); i += 3) {
            const float px = v[i], py = v[i + 1], pz = v[i + 2];
            float out[3];
            for (int r = 0; r < 3; ++r) out[r] = (px * R[0 * 4 + r] + py * R[1 * 4 + r]) + pz * R[2 * 4 + r];
            v[i] = out[0];
            v[i + 1] = out[1];
            v[i + 2] = out[2];
        }
    };
    apply(positions);
    apply(normals);
}

void SceneMesh::setBaseColor(float r, float g, float b) {
    colors.resize(positions.size());
    for (size_t i = 0; i + 2 < colors.size(); i += 3) {
        colors[i] = r;
        colors[i + 1] = g;
        colors[i + 2] = b;
    }
}

int SceneMesh::loadColorFromOBJFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) return -1;
    colors.assign(positions.size(), 0.0f);
    size_t vi = 0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key != "v") continue;
        float f[6] = {0, 0, 0, 0, 0, 0};
        for (float& x : f) {
            std::string tok;
            bool ok = false;
            if (ls >> tok) x = parse_float(tok, ok);
            if (!ok) return -1;  // `v x y z r g b` is required for every vertex
        }
        if (3 * vi + 2 < colors.size()) {
            colors[3 * vi + 0] = f[3];
            colors[3 * vi + 1] = f[4];
            colors[3 * vi + 2] = f[5];
        }
        ++vi;
    }
    return 0;
}

void SceneMesh::loadTexture(const std::string& path, int id) {
    textures.push_back(path);
    ++textureIds;
    for (size_t i = 2; i < textureCoords.size(); i += 3) textureCoords[i] = static_cast<float>(id);
}

void SceneMesh::addObject(const SceneMesh& m) {
    const int32_t base = static_cast<int32_t>(positions.size() / 3);
    positions.insert(positions.end(), m.positions.begin(), m.positions.end());
    normals.insert(normals.end(), m.normals.begin(), m.normals.end());
    textureCoords.insert(textureCoords.end(), m.textureCoords.begin(), m.textureCoords.end());
    // colours concatenate as in the reference: an object without colours adds none
    colors.insert(colors.end(), m.colors.begin(), m.colors.end());
    for (int32_t i : m.indices) indices.push_back(i + base);
    textures.insert(textures.end(), m.textures.begin(), m.textures.end());
    textureIds += m.textureIds;
}

int SceneLoader::load(const std::string& configPath, const std::string& baseDir) {
    std::ifstream in(configPath);
    if (!in) return -1;
    auto resolve_path = [&](const std::string& p) {
        if (p.empty() || p[0] == '/' || baseDir.empty()) return p;
        return baseDir + "/" + p;
    };
    SceneMesh temp;
    bool have = false;
    int numberOfTextures = 0;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream split(line);
        std::string key;
        if (!(split >> key)) continue;
        auto three = [&](float out[3]) {
            for (int a = 0; a < 3; ++a) {
                std::string v;
                split >> v;
                out[a] = static_cast<float>(std::atof(v.c_str()));  // atof -> double -> float
            }
        };
        float v3[3];
        switch (key[0]) {
        case 'o': {
            std::string value;
            split >> value;
            temp = SceneMesh();
            if (temp.loadOBJFile(resolve_path(value)) != 0) return -1;
            temp.computeNormals();
            have = true;
            break;
        }
        case 'm': {
            std::string value;
            split >> value;
            if (!have) return -1;
            ++numberOfTextures;
            temp.loadTexture(value, numberOfTextures);
            break;
        }
        case 's':
            three(v3);
            if (!have) return -1;
            temp.scale(v3[0], v3[1], v3[2]);
            break;
        case 't':
            three(v3);
            if (!have) return -1;
            temp.translate(v3[0], v3[1], v3[2]);
            break;
        case 'r':
            three(v3);
            if (!have) return -1;
            temp.rotate(v3[0], v3[1], v3[2]);
            break;
        case '+':  // addObject, then the reference deletes its temporary mesh
            if (!have) return -1;
            mesh.addObject(temp);
            have = false;
            break;
        case 'v':
            three(v3);
            for (int a = 0; a < 3; ++a) (key.size() > 1 && key[1] == 'e' ? cameraPosition : cameraAt)[a] = v3[a];
            break;
        case 'l':
            three(v3);
            for (int a = 0; a < 3; ++a) (key.size() > 1 && key[1] == 'e' ? lightPosition : lightAt)[a] = v3[a];
            break;
        case 'c':
            if (!have) return -1;
            if (key.size() > 1 && key[1] == 'f') {
                std::string value;
                split >> value;
                if (temp.loadColorFromOBJFile(resolve_path(value)) != 0) return -1;
            } else {
                three(v3);
                temp.setBaseColor(v3[0], v3[1], v3[2]);
            }
            break;
        case 'd': {
            std::string value;
            split >> value;
            depthThreshold = static_cast<float>(std::atof(value.c_str()));
            break;
        }
        case 'h': {  // SoftShadowMapping/src/IO/SceneLoader.cpp:93-96
            std::string value;
            split >> value;
            (key.size() > 1 && key[1] == 'a' ? HSMAlpha : HSMBeta) = static_cast<float>(std::atof(value.c_str()));
            break;
        }
        default:
            break;
        }
    }
    return 0;
}

}  // namespace gi::host

struct gi_scene {
    gi::host::SceneLoader loader;
};

extern "C" {

int gi_scene_load(const char* config_path, const char* base_dir, gi_scene** out) {
    if (!config_path || !out) return GI_EINVAL;
    *out = nullptr;
    gi_scene* s = new (std::nothrow) gi_scene;
    if (!s) return GI_EINVAL;
    if (s->loader.load(config_path, base_dir ? base_dir : "") != 0) {
        delete s;
        return GI_EIO;
    }
    *out = s;
    return GI_OK;
}

int gi_obj_load(const char* obj_path, gi_scene** out) {
    if (!obj_path || !out) return GI_EINVAL;
    *out = nullptr;
    gi_scene* s = new (std::nothrow) gi_scene;
    if (!s) return GI_EINVAL;
    gi::host::SceneMesh m;
    if (m.loadOBJFile(obj_path) != 0) {
        delete s;
        return GI_EIO;
    }
    m.computeNormals();
    s->loader.mesh = std::move(m);
    *out = s;
    return GI_OK;
}

int gi_scene_get_info(const gi_scene* scene, gi_scene_info* out) {
    if (!scene || !out) return GI_EINVAL;
    const auto& L = scene->loader;
    out->nverts = L.mesh.nverts();
    out->ntris = L.mesh.ntris();
    out->has_colors = L.mesh.colors.empty() ? 0 : 1;
    out->ntextures = static_cast<int>(L.mesh.textures.size());
    for (int a = 0; a < 3; ++a) {
        out->camera_eye[a] = L.cameraPosition[a];
        out->camera_at[a] = L.cameraAt[a];
        out->light_eye[a] = L.lightPosition[a];
        out->light_at[a] = L.lightAt[a];
    }
    out->depth_threshold = L.depthThreshold;
    out->hsm_alpha = L.HSMAlpha;
    out->hsm_beta = L.HSMBeta;
    return GI_OK;
}

int gi_scene_copy(const gi_scene* scene, float* positions, float* normals, int32_t* indices, float* colors) {
    if (!scene) return GI_EINVAL;
    const auto& m = scene->loader.mesh;
    if (positions) std::memcpy(positions, m.positions.data(), m.positions.size() * sizeof(float));
    if (normals) std::memcpy(normals, m.normals.data(), m.normals.size() * sizeof(float));
    if (indices) std::memcpy(indices, m.indices.data(), m.indices.size() * sizeof(int32_t));
    if (colors) {
        const size_t n = m.positions.size();
        const size_t c = m.colors.size() < n ? m.colors.size() : n;
        std::memcpy(colors, m.colors.data(), c * sizeof(float));
        std::memset(colors + c, 0, (n - c) * sizeof(float));
    }
    return GI_OK;
}

const char* gi_scene_texture_path(const gi_scene* scene, int index) {
    if (!scene || index < 0 || index >= static_cast<int>(scene->loader.mesh.textures.size())) return nullptr;
    return scene->loader.mesh.textures[static_cast<size_t>(index)].c_str();
}

void gi_scene_free(gi_scene* scene) { delete scene; }

}  // extern "C"
