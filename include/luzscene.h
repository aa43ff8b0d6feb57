/*
 * luzscene.h — C ABI of the host scene model and its GPU packer.
 *
 * The C++17 model behind it keeps Luz's scene classes
 * (source/Resources/AssetManager.hpp): MeshAsset with 48-byte MeshVertex
 * (:82-97), MaterialAsset (:99-112), Node TRS transforms (:114-174,
 * AssetManager.cpp:38-59), MeshNode (:176-182), LightNode (:184-227),
 * CameraNode (:229-275) and SceneAsset settings (:277-296).  The packer
 * mirrors GPUScene (source/Graphics/GPUScene.hpp:35-63): one BLAS per
 * distinct mesh (AddMesh, GPUScene.cpp:108-138), SceneBlock/ModelBlock
 * packing (UpdateResources, :181-346) and the per-frame TLAS rebuild
 * (UpdateResourcesGPU, :348-366).
 *
 * Procedural generators build the measured workloads:
 *   LUZSCENE_CUBE       unit cube on a 20x20 ground quad, one directional
 *                       light, 256x256, 1 shadow + 4 AO rays per pixel;
 *   LUZSCENE_SPHERES    50 instances of 5 fBm-displaced level-5 icospheres
 *                       (20,480 triangles each) over a 64x64 ground grid,
 *                       1 directional + 8 point lights, 1920x1080,
 *                       1 shadow ray per light + 16 AO rays per pixel;
 *   LUZSCENE_TERRAIN    2048x2048 fBm heightfield (decimated 2x) and 40
 *                       displaced icospheres, 4 directional + 32 point
 *                       lights, 3840x2160, 8 AO rays per pixel;
 *   LUZSCENE_RANDOM     the LUZSCENE_SPHERES generator at a seeded triangle
 *                       budget (log-uniform in [1e3, 5e5]) and resolution
 *                       (320x240, 640x480, 1280x720 or 1920x1080): icosphere
 *                       level and count fill the budget, same lights, samples
 *                       and camera (batched workloads).
 */
#ifndef LUZSCENE_H
#define LUZSCENE_H
#include <stdint.h>

#include "luzrt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LUZSCENE_CUBE 1
#define LUZSCENE_SPHERES 2
#define LUZSCENE_TERRAIN 4
#define LUZSCENE_RANDOM 5

typedef struct luzscene luzscene;
typedef struct luzgpu luzgpu;

/* Generate a scene.  For LUZSCENE_RANDOM, tri_budget and width/height are
 * taken from the arguments (0 = draw them from the seed). */
luzscene* luzscene_generate(int kind, uint64_t seed, uint32_t tri_budget, uint32_t width,
                            uint32_t height);
void      luzscene_destroy(luzscene* s);

uint32_t luzscene_mesh_count(const luzscene* s);
/* MeshVertex array (12 floats per vertex: position, normal, tangent, uv) and indices. */
int      luzscene_mesh(const luzscene* s, uint32_t mesh, const float** vertices,
                       uint32_t* vertex_count, const uint32_t** indices, uint32_t* index_count);
uint32_t luzscene_node_count(const luzscene* s);
/* Per mesh node: object->world (row-major 3x4 from Node::GetWorldTransform),
 * mesh index, and the G-buffer encoding of its material (custom index = node). */
int      luzscene_nodes(const luzscene* s, float* object_to_world, uint32_t* mesh_ids,
                        luzrt_material* materials);
/* SceneBlock as GPUScene::UpdateResources packs it (lights, ambient, AO, camera position). */
int      luzscene_pack(const luzscene* s, luzrt_scene* out);
/* Image size and primary-ray camera (CameraNode::GetView/GetProj conventions). */
int      luzscene_camera(const luzscene* s, uint32_t* width, uint32_t* height, luzrt_camera* out);
/* Column-major view-projection (CameraNode::GetProjJittered * GetView) with
 * the given NDC jitter (NULL = none), and its inverse (may be NULL). */
int      luzscene_view_proj(const luzscene* s, const float jitter[2], float view_proj[16],
                            float inv_view_proj[16]);
/* The camera's view and (jittered) projection and their inverses, as
 * SceneBlock carries them (GPUScene.cpp:230-234): column-major, computed in
 * FP64 and rounded once.  inverse_view / inverse_proj are what a depth-mode
 * luzrt_gbuffer needs for DepthToWorld.  Any output may be NULL. */
int      luzscene_camera_matrices(const luzscene* s, const float jitter[2], float view[16], float proj[16],
                                  float inv_view[16], float inv_proj[16]);
/* Shadow-map descriptor of light `light` (GPUScene.cpp:266-312): the
 * orthographic view-projection fitted to the camera frustum for directional
 * and spot lights, the position and shadowMapFar for point lights; depth is
 * left NULL for the caller to allocate (res^2 or 6 res^2 floats). */
int      luzscene_shadow_map(const luzscene* s, uint32_t light, uint32_t resolution, luzrt_shadow_map* out);
/* SceneAsset::shadowType: LUZRT_SHADOW_DISABLED, LUZRT_SHADOW_RAYTRACED or
 * LUZRT_SHADOW_MAP (AssetManager.hpp:40-44). */
int      luzscene_set_shadow_type(luzscene* s, int type);
/* The camera jitter of frame `index` (CameraNode::NextJitter): Halton(2,3). */
void     luzscene_jitter(uint32_t index, uint32_t width, uint32_t height, float jitter[2]);
/* Seeded rigid motion of every mesh node for animation frame `frame`. */
int      luzscene_animate(luzscene* s, uint32_t frame);
/* Total triangles over all mesh nodes (instanced). */
uint64_t luzscene_instanced_triangles(const luzscene* s);

/* Scene ingest (AssetIO.cpp:117-528, AssetManager.cpp:182-257).  The format
 * follows the extension: ".luz" (project JSON; the mesh blob is bin_path, or
 * the .luz path + "bin" when NULL; meshes whose data is absent are skipped
 * and counted in *missing_meshes), ".gltf"/".glb" (glTF 2.0) or ".obj" (with
 * its MTL).  Imported glTF/OBJ scenes get a camera framing their bounds and
 * no lights.  On failure returns NULL and writes a message into error. */
luzscene* luzscene_load(const char* path, const char* bin_path, uint32_t* missing_meshes, char* error,
                        uint32_t error_len);
/* Writes the scene as a .luz project and its .luzbin blob. */
int       luzscene_save(const luzscene* s, const char* path, const char* bin_path, char* error,
                        uint32_t error_len);
/* Camera (eye, center, fov passed as fovy like CameraNode), image size, and
 * near/far distances (0 keeps the current ones). */
int       luzscene_set_camera(luzscene* s, const float eye[3], const float center[3], float fov_degrees,
                              uint32_t width, uint32_t height, float near_distance, float far_distance);
/* Adds a light (LUZRT_LIGHT_*; direction may be NULL for point lights). */
int       luzscene_add_light(luzscene* s, int type, const float position[3], const float direction[3],
                             const float color[3], float intensity, float radius);
/* SceneAsset::lightSamples / aoSamples / aoMin / aoMax. */
int       luzscene_set_samples(luzscene* s, int light_samples, int ao_samples, float ao_min, float ao_max);

/* GPUScene: uploads meshes, builds one BLAS per mesh and the TLAS. */
luzgpu*  luzgpu_create(luzrt_ctx* ctx, const luzscene* s);
void     luzgpu_destroy(luzgpu* g);
/* Re-pack instances from the scene and rebuild the TLAS (per-frame work). */
luzrt_status luzgpu_update(luzgpu* g, const luzscene* s);
luzrt_tlas*  luzgpu_tlas(luzgpu* g);
luzrt_blas*  luzgpu_blas(luzgpu* g, uint32_t mesh);
/* Device vertex buffer (MeshVertex array) of a mesh, as its BLAS was built from. */
const void*  luzgpu_vertices(luzgpu* g, uint32_t mesh);
/* The TLAS instances of the scene's mesh nodes (out may be NULL); returns the count. */
uint32_t     luzgpu_instances(luzgpu* g, const luzscene* s, luzrt_instance* out);

#ifdef __cplusplus
}
#endif
#endif
