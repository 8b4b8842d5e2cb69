/* gi_scene.h — scene configuration and OBJ loading for the shadow pipeline.
 *
 * Replaces the reference apps' scene input path, which the passes consume:
 *   SceneLoader::load      ShadowMapping/src/IO/SceneLoader.cpp:12-91
 *                          (SoftShadowMapping/src/IO/SceneLoader.cpp adds `ha`/`hb`)
 *   Mesh::loadOBJFile      ShadowMapping/src/Mesh.cpp:251-313 over glmReadOBJ
 *                          (ShadowMapping/src/IO/OBJLoader.cpp:581-770)
 *   Mesh::computeNormals   Mesh.cpp:195-234
 *   Mesh::scale/translate/rotate/setBaseColor/loadColorFromOBJFile  Mesh.cpp:327-427
 *   Mesh::addObject        Mesh.cpp:47-192
 * Host code only: nothing here touches the GPU.  Every entry point returns a
 * gi_status (include/gi.h); GI_EIO reports a file that cannot be opened.
 */
#ifndef GI_SCENE_H
#define GI_SCENE_H

#include <stddef.h>
#include <stdint.h>

#include "gi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gi_scene gi_scene;

typedef struct gi_scene_info {
    int nverts;
    int ntris;
    int has_colors;          /* per-vertex RGB from `c r g b` or `cf file` */
    int ntextures;           /* `m` lines (image paths; images are not decoded) */
    float camera_eye[3];     /* `ve` (SceneLoader cameraPosition) */
    float camera_at[3];      /* `va` */
    float light_eye[3];      /* `le` */
    float light_at[3];       /* `la` */
    float depth_threshold;   /* `d` */
    float hsm_alpha;         /* `ha` (soft app) */
    float hsm_beta;          /* `hb` (soft app) */
} gi_scene_info;

/* Loads a scene configuration (Configs/<name>.txt).  Paths inside it are
 * resolved against base_dir (the reference runs with the repository root as
 * its working directory); base_dir may be NULL for the current directory. */
int gi_scene_load(const char* config_path, const char* base_dir, gi_scene** out);

/* One OBJ file: Mesh::loadOBJFile followed by Mesh::computeNormals, as the
 * loader's `o` line does. */
int gi_obj_load(const char* obj_path, gi_scene** out);

int gi_scene_get_info(const gi_scene* scene, gi_scene_info* out);

/* Copies the merged mesh: positions/normals nverts*3 floats, indices ntris*3
 * int32, colors nverts*3 floats (may be NULL; zeros when has_colors is 0). */
int gi_scene_copy(const gi_scene* scene, float* positions, float* normals, int32_t* indices, float* colors);

/* Path of texture `index` (0-based) as written in the configuration. */
const char* gi_scene_texture_path(const gi_scene* scene, int index);

void gi_scene_free(gi_scene* scene);

#ifdef __cplusplus
}
#endif

#endif /* GI_SCENE_H */
