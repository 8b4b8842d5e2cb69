// scene_loader.hpp — scene configuration and OBJ input of the shadow pipeline.
//
// The reference builds its scene with SceneLoader::load
// (ShadowMapping/src/IO/SceneLoader.cpp:12-91) over Mesh
// (ShadowMapping/src/Mesh.cpp): each `o` line reads an OBJ file through
// glmReadOBJ (ShadowMapping/src/IO/OBJLoader.cpp), recomputes per-vertex
// normals, applies the `s`/`t`/`r` lines in file order and is appended to the
// scene mesh at `+`.  These classes keep that order and those numerics (FP32
// geometry, the FP64 running mean of face normals) so the meshes handed to the
// passes are the reference's.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace gi::host {

// The Mesh fields the passes and the G-buffer colour read.
struct SceneMesh {
    std::vector<float> positions;      // nverts x 3
    std::vector<float> normals;        // nverts x 3
    std::vector<float> textureCoords;  // nverts x 3 (u, v, texture id), Mesh.cpp:287-301
    std::vector<float> colors;         // nverts x 3 or empty
    std::vector<int32_t> indices;      // ntris x 3, 0-based
    std::vector<std::string> textures; // `m` paths, in order
    int textureIds = 0;                // Mesh::numberOfTextures

    int nverts() const { return static_cast<int>(positions.size() / 3); }
    int ntris() const { return static_cast<int>(indices.size() / 3); }

    int loadOBJFile(const std::string& path);             // Mesh.cpp:251-313
    void computeNormals();                                 // Mesh.cpp:195-234
    void scale(float x, float y, float z);                 // Mesh.cpp:390-400
    void translate(float x, float y, float z);             // Mesh.cpp:378-388
    void rotate(float x, float y, float z);                // Mesh.cpp:402-427 (degrees)
    void setBaseColor(float r, float g, float b);          // Mesh.cpp:365-376
    int loadColorFromOBJFile(const std::string& path);     // Mesh.cpp:327-363
    void loadTexture(const std::string& path, int id);     // Mesh.cpp:315-325 (path only)
    void addObject(const SceneMesh& m);                    // Mesh.cpp:47-192
};

class SceneLoader {
public:
    SceneMesh mesh;
    std::array<float, 3> cameraPosition{}, cameraAt{}, lightPosition{}, lightAt{};
    float depthThreshold = 0.0f;
    float HSMAlpha = 0.0f, HSMBeta = 0.0f;

    // baseDir: the directory the configuration's relative paths start from.
    int load(const std::string& configPath, const std::string& baseDir);
};

}  // namespace gi::host
