/*
 * luz_oracle.h — CPU oracle for the Luz ray-traced lighting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry point and the CPU-baseline leg of bench.py load it to compare the HIP
 * path against.  The product library (libluzrt) never links, loads or calls it.
 *
 * It restates, in plain FP32/FP64 C++ on the CPU:
 *   - the acceleration-structure build that Luz delegates to the Vulkan driver
 *     (VulkanWrapper.cpp:743-836 CreateBLAS, :1089-1104 CmdBuildBLAS,
 *      :1106-1139 CmdBuildTLAS) as a deterministic LBVH (Karras 2012);
 *   - the ray-query visibility semantics of light.frag:86-135 (any-hit,
 *     terminate on first hit, opaque, cull disabled) with the
 *     Woop-Benthin-Wald 2013 watertight triangle test;
 *   - the shadow/AO ray construction of light.frag:86-146,229-231;
 *   - the GGX shading of light.frag:17-49,171-235;
 *   - a primary-ray G-buffer producer standing in for opaque.frag:21-58.
 *
 * Parity status of each function is recorded in the header of luz_oracle.cpp.
 */
#ifndef LUZ_ORACLE_H
#define LUZ_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

/* LightBlock subset (LuzCommon.h:37-61); same 64-byte layout as luzrt_light. */
typedef struct orc_light {
    float color[3];     float intensity;
    float position[3];  float inner_angle;
    float direction[3]; float outer_angle;
    int32_t type;       int32_t num_shadow_samples;
    float radius;       int32_t pad;
} orc_light;

/* SceneBlock subset (LuzCommon.h:94-125); same layout as luzrt_scene. */
typedef struct orc_params {
    orc_light lights[64];
    float ambient_color[3]; float ambient_intensity;
    float cam_pos[3];       int32_t num_lights;
    float ao_min;           float ao_max;
    int32_t ao_num_samples; int32_t frame;
    int32_t shadow_type;    int32_t pad[3];
} orc_params;

typedef struct orc_gbuffer {
    uint32_t width, height;
    const float* position;     /* float3 per pixel */
    const float* normal;       /* float3 per pixel; (0,0,0) = background */
    const uint32_t* albedo;    /* RGBA8 */
    const uint32_t* material;  /* RGBA8: r roughness, g metallic, b occlusion */
    const uint32_t* emission;  /* RGBA8 */
} orc_gbuffer;

typedef struct orc_camera {
    float eye[3];
    float forward[3];   /* ray direction through the image centre */
    float right[3];     /* added per unit of NDC x */
    float up[3];        /* added per unit of NDC y (y points down the image) */
    float tmin, tmax;
} orc_camera;

typedef struct orc_material {  /* per custom index, RGBA8 each */
    uint32_t albedo, material, emission, pad;
} orc_material;

typedef struct orc_stats {
    uint64_t rays, node_visits, tri_tests, inst_visits, hits;
    uint64_t wide_node_visits;   /* 96-byte nodes of the 8-wide BLASes */
} orc_stats;

/* --- primitives exposed for known-answer tests ------------------------- */
uint32_t orc_pcg(uint32_t v);
void     orc_pcg4d(const uint32_t in[4], uint32_t out[4]);
void     orc_sample2(uint32_t x, uint32_t y, uint32_t frame, uint32_t slot, float uv[2]);
void     orc_sincos_turns(float w, float* s, float* c);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);
/* WBW watertight test: returns 1 on hit with t in [tmin,tmax]; writes t,bary */
int      orc_intersect_tri(const float o[3], const float d[3], float tmin, float tmax,
                           const float v0[3], const float v1[3], const float v2[3],
                           float* t_out, float bary_out[3]);
int      orc_box_test(const float o[3], const float d[3], float tmin, float tmax,
                      const float lo[3], const float hi[3]);
void     orc_invert_3x4(const float m[12], float out[12]);

/* --- scene / acceleration structures ----------------------------------- */
orc_scene* orc_scene_create(void);
void       orc_scene_destroy(orc_scene* s);
/* vertices: positions at byte offset 0 (normals at 12 when stride >= 24). */
int  orc_scene_add_mesh(orc_scene* s, const void* vertices, uint32_t stride_bytes,
                        uint32_t vertex_count, const uint32_t* indices, uint32_t index_count);
/* Replace the vertex positions of a mesh and refit its BLAS (topology kept). */
int  orc_scene_refit_mesh(orc_scene* s, int mesh, const void* vertices);
/* object_to_world: n * 12 floats, row-major 3x4. */
int  orc_scene_set_instances(orc_scene* s, const float* object_to_world,
                             const uint32_t* mesh_ids, const uint32_t* custom_index,
                             const uint32_t* mask, uint32_t n);
int  orc_blas_arrays(const orc_scene* s, int mesh, const void** nodes, uint32_t* node_count,
                     const float** tris, const uint32_t** prim_ids, uint32_t* tri_count);
int  orc_tlas_arrays(const orc_scene* s, const void** nodes, uint32_t* node_count,
                     const uint32_t** inst_order, const float** world_to_object,
                     uint32_t* inst_count);

/* rays: n * 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax) in world space */
int  orc_any_hit(const orc_scene* s, const float* rays, uint32_t n, uint8_t* hits,
                 int brute_force, orc_stats* stats);
/* closest hit: t (inf on miss), instance index (input order), primitive id */
int  orc_closest_hit(const orc_scene* s, const float* rays, uint32_t n, float* t,
                     uint32_t* inst, uint32_t* prim, int brute_force);

/* --- the lighting path --------------------------------------------------- */
int  orc_primary_gbuffer(const orc_scene* s, const orc_camera* cam, const orc_material* mats,
                         uint32_t num_mats, uint32_t width, uint32_t height,
                         float* position, float* normal, uint32_t* albedo,
                         uint32_t* material, uint32_t* emission, int threads);
/* shadow: [num_lights][W*H] occluded-sample counts; ao: [W*H] unoccluded counts.
 * Only rows 0, row_step, 2*row_step, ... are traced (1 = every row); other
 * rows of the outputs are left untouched. */
int  orc_trace_visibility(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                          uint8_t* shadow, uint8_t* ao, orc_stats* stats, int threads,
                          uint32_t row_step);
/* Same rays in FP64 (ray set-up, sampling angles and intersection). */
int  orc_depth_to_world(const float* depth, uint32_t W, uint32_t H, uint32_t pitch, const float inv_proj[16],
                        const float inv_view[16], float* position);
int  orc_any_hit_f64(const orc_scene* s, const double* rays, uint32_t n, uint8_t* hits);
int  orc_trace_visibility_f64(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                              uint8_t* shadow, uint8_t* ao, int threads, uint32_t row_step);
/* Precision form: 8-wide BLASes with 8-bit quantised child boxes, built from
 * each mesh's binary LBVH; visibility traced through them (TLAS binary). */
int  orc_scene_build_wide(orc_scene* s);
int  orc_blas_wide_arrays(const orc_scene* s, int mesh, const void** nodes, uint32_t* node_count,
                          const float** tris, const uint32_t** prim_ids, uint32_t* tri_count);
int  orc_any_hit_wide(const orc_scene* s, const float* rays, uint32_t n, uint8_t* hits,
                      orc_stats* stats);
int  orc_trace_visibility_wide(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                               uint8_t* shadow, uint8_t* ao, orc_stats* stats, int threads,
                               uint32_t row_step);
/* Emit the ray list the visibility pass casts (for brute-force checks). */
int64_t orc_visibility_rays(const orc_gbuffer* g, const orc_params* p, float* rays,
                            int64_t capacity);
int  orc_shade(const orc_gbuffer* g, const orc_params* p, const uint8_t* shadow,
               const uint8_t* ao, float* out_rgba, int threads);
int  orc_shade_f64(const orc_gbuffer* g, const orc_params* p, const uint8_t* shadow,
                   const uint8_t* ao, double* out_rgba, int threads);

/* ---- passes after the light pass (post_oracle.cpp) ---- */
typedef struct orc_taa_desc {
    uint32_t width, height;
    const float* light;      /* RGBA32F */
    const float* history;    /* RGBA32F */
    const float* position;   /* float3 */
    const float* normal;     /* float3, zero = background */
    float view_proj[16], prev_view_proj[16], inv_view_proj[16];   /* column-major */
    float jitter[2], prev_jitter[2];
    int32_t reconstruct;
} orc_taa_desc;

int orc_taa(const orc_taa_desc* d, float* out_rgba32f);
int orc_compose(const float* light_rgba32f, uint32_t npix, uint32_t* rgba8, float* rgba32f);

#ifdef __cplusplus
}
#endif
#endif
