/*
 * luzrt.h — C ABI of the MI355X-native ray-traced lighting path.
 *
 * This replaces three interfaces of Luz (hadryansalles/Luz):
 *   1. the acceleration-structure part of the Vulkan wrapper
 *        vkw::CreateBLAS / CmdBuildBLAS   source/Graphics/VulkanWrapper.h:202-209,242
 *                                         source/Graphics/VulkanWrapper.cpp:743-836,1089-1104
 *        vkw::CreateTLAS / CmdBuildTLAS   source/Graphics/VulkanWrapper.h:192-200,243
 *                                         source/Graphics/VulkanWrapper.cpp:657-741,1106-1139
 *   2. the light pass  DeferredRenderer::LightPass(LightConstants)
 *                                         source/Graphics/DeferredRenderer.hpp:17-26,46
 *                                         source/Graphics/DeferredRenderer.cpp:324-345
 *      together with the device program it launches, source/Shaders/light.frag:86-235;
 *   3. the host<->device blocks it reads   source/Shaders/LuzCommon.h:37-61 (LightBlock),
 *                                         :94-125 (SceneBlock), :180-189 (LightConstants).
 *
 * Conventions
 *   - Every function returns a luzrt_status; no exception crosses the ABI.
 *     luzrt_last_error() returns a message for the last failure on a context.
 *   - Handles are opaque and owned by the library; destroy them explicitly.
 *     A BLAS referenced by a TLAS that still exists cannot be destroyed
 *     (LUZRT_E_IN_USE): the reference keeps each BLAS alive through GPUMesh
 *     (GPUScene.hpp:17-23) because its TLAS instances hold the BLAS address.
 *   - Buffers named "device" are device pointers (luzrt_malloc, hipMalloc or a
 *     torch tensor's data_ptr) on the context's device; "host" are host memory.
 *   - Work is ordered on one HIP stream per context and is asynchronous until
 *     luzrt_sync().  A context must not be used from two threads at once.
 *   - Limits follow LuzCommon.h:20-21: 64 lights, 8192 instances.  Beyond
 *     them calls fail with LUZRT_E_LIMIT (the reference overflows silently,
 *     GPUScene.cpp:51,240).
 */
#ifndef LUZRT_H
#define LUZRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUZRT_MAX_LIGHTS 64       /* LuzCommon.h:20 LUZ_MAX_LIGHTS */
#define LUZRT_MAX_INSTANCES 8192  /* LuzCommon.h:21 LUZ_MAX_MODELS */
#define LUZRT_MAX_SAMPLES 255     /* per-pixel counts are stored in 8 bits */
#define LUZRT_MAX_BATCH 256       /* scenes per batched frame */

#define LUZRT_LIGHT_POINT 0       /* LuzCommon.h:24-26 */
#define LUZRT_LIGHT_SPOT 1
#define LUZRT_LIGHT_DIRECTIONAL 2

#define LUZRT_SHADOW_DISABLED 0   /* AssetManager.hpp:41 ShadowDisabled */
#define LUZRT_SHADOW_RAYTRACED 1  /* LuzCommon.h:28 SHADOW_TYPE_RAYTRACING */
#define LUZRT_SHADOW_MAP 2        /* LuzCommon.h:29 SHADOW_TYPE_MAP */

typedef enum luzrt_status {
    LUZRT_OK = 0,
    LUZRT_E_INVALID_ARG = -1,
    LUZRT_E_OUT_OF_MEMORY = -2,
    LUZRT_E_DEVICE = -3,      /* a HIP call failed */
    LUZRT_E_IN_USE = -4,      /* BLAS still referenced by a TLAS */
    LUZRT_E_LIMIT = -5,       /* lights, instances, samples or sizes over the limits */
    LUZRT_E_NO_DEVICE = -6,   /* no HIP device / bad device index */
    LUZRT_E_UNSUPPORTED = -7
} luzrt_status;

typedef struct luzrt_ctx luzrt_ctx;
typedef struct luzrt_blas luzrt_blas;
typedef struct luzrt_tlas luzrt_tlas;

/* BLASInstance (VulkanWrapper.h:196-200) as packed by CmdBuildTLAS
 * (VulkanWrapper.cpp:1113-1126): row r of object_to_world is row r of the
 * column-major modelMat, custom_index is the ModelBlock index, mask 0xFF
 * for a visible instance (0 removes it from every ray). */
typedef struct luzrt_instance {
    const luzrt_blas* blas;
    float object_to_world[3][4];
    uint32_t custom_index;
    uint32_t mask;
} luzrt_instance;

/* LightBlock (LuzCommon.h:37-61), the part the lighting path reads.
 * direction is the light node's W*(0,-1,0,0) and need not be unit length
 * (GPUScene.cpp:245); angles are in radians (GPUScene.cpp:244,246). */
typedef struct luzrt_light {
    float color[3];     float intensity;
    float position[3];  float inner_angle;
    float direction[3]; float outer_angle;
    int32_t type;       int32_t num_shadow_samples;
    float radius;       int32_t pad;
} luzrt_light;

/* SceneBlock (LuzCommon.h:94-125), the part the lighting path reads, plus
 * LightConstants::frame (LuzCommon.h:183). */
typedef struct luzrt_scene {
    luzrt_light lights[LUZRT_MAX_LIGHTS];
    float ambient_color[3]; float ambient_intensity;
    float cam_pos[3];       int32_t num_lights;
    float ao_min;           float ao_max;
    int32_t ao_num_samples; int32_t frame;
    int32_t shadow_type;    int32_t pad[3];
} luzrt_scene;

/* The G-buffer of DeferredRenderer::CreateImages (DeferredRenderer.cpp:175-248),
 * device pointers.  A zero normal marks a background pixel (light.frag:178).
 *
 * Position comes either from a world-space position plane or, with
 * position == NULL, from the engine's D32F depth attachment
 * (DeferredRenderer.cpp:232-238) through DepthToWorld (utils.glsl:1-7):
 * clip = (uv * 2 - 1, depth, 1) at the pixel centre uv = ((x + 0.5) / width,
 * (y + 0.5) / height), v = inverse_proj * clip, v /= v.w, P = (inverse_view *
 * v).xyz, with SceneBlock's inverseProj / inverseView (LuzCommon.h:104-105,
 * GPUScene.cpp:233-234), column-major, products summed over the columns in
 * order and the divisions correctly rounded.
 *
 * Every plane is read in place with a row pitch and (position and normal)
 * a pixel stride in bytes, so the engine's images are passed as they are:
 * normal_stride 16 reads the RGB of the RGBA32F normal attachment
 * (DeferredRenderer.cpp:190-196).  Zero layout fields (a zero-initialised
 * tail) mean tightly packed rows: stride 12, pitch width * stride (width * 4
 * for the RGBA8 planes and the depth). */
typedef struct luzrt_gbuffer {
    uint32_t width, height;
    const float* position;     /* float3 per pixel, or NULL: positions from depth */
    const float* normal;       /* float3 per pixel (first three floats of normal_stride bytes) */
    const uint32_t* albedo;    /* RGBA8 sRGB (light.frag:172 decodes with pow 2.2) */
    const uint32_t* material;  /* RGBA8: r roughness, g metallic, b occlusion (opaque.frag:57) */
    const uint32_t* emission;  /* RGBA8 */
    /* layout: pixel strides (12 or 16) and row pitches in bytes, 0 = packed */
    uint32_t position_stride, normal_stride;
    uint32_t position_pitch, normal_pitch, albedo_pitch, material_pitch, emission_pitch;
    /* depth mode (position == NULL): D32F depth, its row pitch (0 = width * 4) */
    uint32_t depth_pitch;
    const float* depth;
    float inverse_proj[16];
    float inverse_view[16];
} luzrt_gbuffer;

/* Writable G-buffer for the primary-ray producer. */
typedef struct luzrt_gbuffer_out {
    uint32_t width, height;
    float* position; float* normal;
    uint32_t* albedo; uint32_t* material; uint32_t* emission;
} luzrt_gbuffer_out;

/* Pinhole camera for primary rays: direction(x,y) = forward + ndc_x*right +
 * ndc_y*up with ndc = (pixel centre / size)*2 - 1 and ndc_y growing down the
 * image (the y flip of CameraNode::GetProj, AssetManager.cpp:375). */
typedef struct luzrt_camera {
    float eye[3]; float forward[3]; float right[3]; float up[3];
    float tmin, tmax;
} luzrt_camera;

/* ModelBlock material fields (LuzCommon.h:76-92) in G-buffer encoding,
 * indexed by luzrt_instance::custom_index. */
typedef struct luzrt_material {
    uint32_t albedo, material, emission, pad;
} luzrt_material;

/* GPU time of the stages of the last call, from HIP events on the
 * context's stream (the reference's CmdTimeStamp, VulkanWrapper.cpp:1276-1296).
 * Visibility calls generate their rays inside the traversal: trace_ms
 * includes ray generation and raygen_ms covers only the ray count. */
typedef struct luzrt_timings {
    float build_ms, compact_ms, raygen_ms, trace_ms, resolve_ms, shade_ms, primary_ms, denoise_ms,
          post_ms;
    uint64_t rays;          /* rays cast by the last visibility/light call */
    uint64_t active_pixels; /* non-background pixels of that call */
    uint64_t traced_rays;   /* of those, ray ids the traversal ran (64 per live row when the
                               call culled ray groups, luzrt_set_group_culling; else = rays) */
} luzrt_timings;

/* ------------------------------------------------------------ context */
luzrt_status luzrt_create(luzrt_ctx** out, int device);
luzrt_status luzrt_destroy(luzrt_ctx* ctx);
const char*  luzrt_last_error(const luzrt_ctx* ctx);
const char*  luzrt_status_string(luzrt_status s);
/* Waits for the context's work.  Returns LUZRT_E_LIMIT (once; the condition
 * is then cleared) when a visibility, light-pass, ray, primary or shadow-map
 * call since the last luzrt_sync / luzrt_get_timings had a ray that exceeded
 * the traversal stack or step limit: such a ray was written as a miss (the
 * device latches the condition; no later call clears it).  The reference has
 * no such limit (the driver's traversal) and no error path (Base.hpp:51-59). */
luzrt_status luzrt_sync(luzrt_ctx* ctx);
/* Per-stage GPU times of the last build/trace/light/primary call (synchronises);
 * 0 for stages run while stage timing is off. */
luzrt_status luzrt_get_timings(luzrt_ctx* ctx, luzrt_timings* out);
/* Stage timing on (the default) or off.  On, every stage records a HIP event
 * at each end, which holds the next kernel back by about 8 us; frames
 * submitted back to back run without them when it is off.  The reference
 * always writes its timestamps (CmdBeginTimeStamp / CmdEndTimeStamp,
 * VulkanWrapper.h:250-267, VulkanWrapper.cpp:1276-1296), which cost its
 * queue nothing comparable. */
luzrt_status luzrt_set_stage_timing(luzrt_ctx* ctx, int enable);
/* The context's device: compute units and peak engine clock (kHz), as the
 * HIP runtime reports them (for throughput and issue-rate figures). */
luzrt_status luzrt_device_info(const luzrt_ctx* ctx, uint32_t* compute_units, uint32_t* clock_khz);
/* Traversal launch parameters (0 keeps the current value): persistent
 * workgroups per CU (wave occupancy), idle lanes of a wave that trigger a
 * refill, LDS stack entries per lane (8, 16 or 32; deeper entries spill to
 * global memory) and the width of the 64-pixel ray tiles (8, 16 or 32).
 * Chosen by the autotuner (luz_amd/autotune.py). */
luzrt_status luzrt_set_trace_config(luzrt_ctx* ctx, uint32_t groups_per_cu, uint32_t refill_threshold,
                                    uint32_t stack_entries, uint32_t tile_width);

/* Traversal steps (node visits and instance entries) a ray may take before
 * it is abandoned as a miss and the error latch reported by luzrt_sync is
 * set; 0 restores the default 2^24, which a well-formed structure never
 * reaches (a guard against malformed trees).  Visibility and ray calls. */
luzrt_status luzrt_set_step_limit(luzrt_ctx* ctx, uint32_t max_steps);

/* Order of a visibility call's rays (a traversal launch parameter, like the
 * above; no counterpart in the reference, whose ray queries run per
 * fragment).  SLOT: all pixels' first sample, then all second samples, ...;
 * GROUPED: the samples of 64 neighbouring pixels together, one sample after
 * the other; AUTO (default): GROUPED when the call can cull its ray groups
 * (luzrt_set_group_culling on, at most 128 TLAS instances and 64 rays per
 * pixel) or while the distinct BLASes of the TLAS total at most 32 MB, SLOT
 * otherwise.  Results are the same in every order.
 * Batched frames use SLOT unless GROUPED is set before luzrt_batch_create. */
enum { LUZRT_RAY_ORDER_AUTO = 0, LUZRT_RAY_ORDER_SLOT = 1, LUZRT_RAY_ORDER_GROUPED = 2 };
luzrt_status luzrt_set_ray_order(luzrt_ctx* ctx, uint32_t order);

/* Ray-group culling of visibility and light-pass calls (a launch parameter,
 * like the above; no counterpart in the reference).  Enabled (default): when
 * the rays are in GROUPED order, the TLAS holds at most 128 instances and a
 * pixel casts at most 64 rays, the rays of each 64-pixel group and light (or
 * the group's AO rays) are bounded by a region computed from the group's
 * G-buffer texels; instances whose boxes miss that region are not visited,
 * and a group and light with no such instance has its rays written as misses
 * without tracing them.  The bounds are conservative, so the results are the
 * same as without culling.  luzrt_get_timings reports the rays traced. */
luzrt_status luzrt_set_group_culling(luzrt_ctx* ctx, int enable);

/* Precision modes.  PLAIN: binary BVH with FP32 boxes, FP32 shading.
 * COMPACT: BLASes traversed as 8-wide nodes with 8-bit child boxes rounded
 * outward (96-byte nodes), FP32 watertight triangle tests, GGX shading with
 * the Fresnel and diffuse terms in packed FP16 (two lights per half2), the
 * specular scalar D*G/denominator in FP32, FP32 accumulation.
 * WIDE: COMPACT's 8-wide BVH with PLAIN's FP32 shading.  Every box test is
 * conservative (the far slab bound is widened past its rounding error) and
 * the triangle tests are the same FP32 watertight tests, so a ray meets the
 * same triangles through either tree: counts and radiance are PLAIN's, bit
 * for bit, from fewer dependent node reads.  Batched frames are PLAIN only.
 * A mode takes effect at the next TLAS build. */
#define LUZRT_PRECISION_PLAIN 0
#define LUZRT_PRECISION_COMPACT 1
#define LUZRT_PRECISION_WIDE 2
luzrt_status luzrt_set_precision(luzrt_ctx* ctx, int mode);

/* ------------------------------------------------------------ memory */
luzrt_status luzrt_malloc(luzrt_ctx* ctx, size_t bytes, void** device_ptr);
luzrt_status luzrt_free(luzrt_ctx* ctx, void* device_ptr);
luzrt_status luzrt_upload(luzrt_ctx* ctx, void* device_dst, const void* host_src, size_t bytes);
luzrt_status luzrt_download(luzrt_ctx* ctx, void* host_dst, const void* device_src, size_t bytes);

/* ------------------------------------------------------ acceleration structures */
/* CreateBLAS + CmdBuildBLAS.  vertices: device, positions (float3) at byte
 * offset 0 of each stride-sized vertex, normals at offset 12 when stride >= 24
 * (MeshVertex, AssetManager.hpp:83-89); indices: device u32, index_count/3
 * triangles.  The caller's buffers are read during the call only. */
luzrt_status luzrt_blas_build(luzrt_ctx* ctx, const void* vertices, uint32_t stride_bytes,
                              uint32_t vertex_count, const uint32_t* indices,
                              uint32_t index_count, luzrt_blas** out);
/* Recompute the boxes of an existing BLAS from moved vertices (same topology). */
luzrt_status luzrt_blas_refit(luzrt_ctx* ctx, luzrt_blas* blas, const void* vertices);
luzrt_status luzrt_blas_destroy(luzrt_ctx* ctx, luzrt_blas* blas);
/* Device pointers of the built structure (64-byte nodes, 36-byte triangles,
 * slot -> primitive ids) for inspection. */
luzrt_status luzrt_blas_arrays(const luzrt_blas* blas, const void** nodes, uint32_t* node_count,
                               const float** tris, const uint32_t** prim_ids, uint32_t* tri_count);

/* Device pointers of the 8-wide form (96-byte nodes, 36-byte triangles in
 * wide-leaf order, primitive ids); built on first use in COMPACT mode. */
luzrt_status luzrt_blas_wide_arrays(luzrt_ctx* ctx, luzrt_blas* blas, const void** nodes,
                                    uint32_t* node_count, const float** tris,
                                    const uint32_t** prim_ids, uint32_t* tri_count);

/* CreateTLAS(maxInstances). */
luzrt_status luzrt_tlas_create(luzrt_ctx* ctx, uint32_t max_instances, luzrt_tlas** out);
/* CmdBuildTLAS: full rebuild from host instances (VulkanWrapper.cpp:1106-1139). */
luzrt_status luzrt_tlas_build(luzrt_ctx* ctx, luzrt_tlas* tlas, const luzrt_instance* instances,
                              uint32_t count);
luzrt_status luzrt_tlas_destroy(luzrt_ctx* ctx, luzrt_tlas* tlas);
/* Device pointers: nodes, world->object 3x4 per TLAS slot, slot -> input index. */
luzrt_status luzrt_tlas_arrays(const luzrt_tlas* tlas, const void** nodes, uint32_t* node_count,
                               const float** world_to_object, const uint32_t** inst_order,
                               uint32_t* inst_count);

/* ------------------------------------------------------------ ray queries */
/* Any-hit queries (gl_RayFlagsTerminateOnFirstHitEXT, mask 0xFF; light.frag:100,126)
 * over device rays, 8 floats each: origin.xyz, tmin, direction.xyz, tmax.
 * hits: device u8 per ray. */
luzrt_status luzrt_trace_rays(luzrt_ctx* ctx, const luzrt_tlas* tlas, const float* rays,
                              uint32_t count, uint8_t* hits);

/* ------------------------------------------------------------ the lighting path */
/* Shadow + AO visibility for every G-buffer pixel (light.frag:86-146,229-231).
 * shadow_counts: device u8 [num_lights][width*height], occluded samples per light;
 * ao_counts: device u8 [width*height], unoccluded AO samples. */
luzrt_status luzrt_trace_visibility(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                    const luzrt_gbuffer* gbuffer, const luzrt_scene* scene,
                                    uint8_t* shadow_counts, uint8_t* ao_counts);
/* luzrt_trace_visibility without any allocation on the call path: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc).  Its work buffers live in one device array of 60.1 MiB
 * that the library holds on each GPU it runs on until the process ends.
 * Limits: width*height <= 2^19 pixels and width*height x (shadow samples of
 * all lights + AO samples) <= 2^24 rays (LUZRT_E_LIMIT otherwise); at most
 * 768 persistent traversal workgroups; a ray's stack holds 48 entries (a
 * deeper one is reported like any traversal overflow, by luzrt_sync).
 * Results equal luzrt_trace_visibility's.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (two contexts' streams) race. */
luzrt_status luzrt_trace_visibility_static(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                           const luzrt_gbuffer* gbuffer, const luzrt_scene* scene,
                                           uint8_t* shadow_counts, uint8_t* ao_counts);
/* GGX lighting from a G-buffer and visibility counts (light.frag:171-235). */
luzrt_status luzrt_shade(luzrt_ctx* ctx, const luzrt_gbuffer* gbuffer, const luzrt_scene* scene,
                         const uint8_t* shadow_counts, const uint8_t* ao_counts,
                         float* rgba32f_out);
/* DeferredRenderer::LightPass: G-buffer in, RGBA32F light buffer out
 * (device, width*height*4 floats). */
/* The lighting pass from visibility fractions instead of counts: visibility
 * [num_lights][npix] (1 = unshadowed) and AO [npix] fractions in [0, 1], such
 * as the learned denoiser's outputs.  Same arithmetic as luzrt_shade with
 * shadow = 1 - visibility and AO as given. */
luzrt_status luzrt_shade_fractions(luzrt_ctx* ctx, const luzrt_gbuffer* g, const luzrt_scene* scene,
                                   const float* visibility, const float* ao, float* rgba32f_out);
luzrt_status luzrt_light_pass(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* gbuffer,
                              const luzrt_scene* scene, float* rgba32f_out);

/* Primary-ray G-buffer producer (closest hit; stands in for the raster
 * opaque pass, opaque.frag:21-58).  materials: host array indexed by
 * custom_index; unknown indices get white, roughness 0.5, metallic 0. */
luzrt_status luzrt_primary_gbuffer(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                   const luzrt_camera* camera, const luzrt_material* materials,
                                   uint32_t material_count, const luzrt_gbuffer_out* out);

/* ------------------------------------------------------------ shadow maps */
/* The alternative visibility source of EvaluateShadow (light.frag:147-165,
 * shadowType == LUZRT_SHADOW_MAP).  A directional or spot light's map is
 * the depth (NDC z of an orthographic view-projection fitted to the camera
 * frustum, GPUScene.cpp:278-311) of the nearest surface through each texel
 * centre; a point light's is a cube map (faces +X -X +Y -Y +Z -Z, Vulkan
 * cube-face conventions) of distance / z_far (shadowMap.frag:11-19).  Maps
 * are rendered here by closest-hit rays through texel centres (the
 * reference rasterises them), cleared to 1.  Lookups sample the map
 * bilinearly: REPEAT addressing for 2D maps (VulkanWrapper.cpp:2433-2437),
 * clamped within the face for cube maps. */
typedef struct luzrt_shadow_map {
    int32_t light_type;         /* LUZRT_LIGHT_* of the light */
    uint32_t resolution;        /* texels per side */
    float view_proj[16];        /* directional/spot: column-major light view-projection */
    float inv_view_proj[16];    /* its inverse */
    float light_pos[3];         /* point lights */
    float z_far;                /* point lights: LightNode::shadowMapFar */
    float* depth;               /* device: res*res floats, or 6*res*res for point lights */
} luzrt_shadow_map;

/* Renders map->depth (allocated by the caller) over the TLAS. */
luzrt_status luzrt_shadow_map_build(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_shadow_map* map);
/* Shadow decisions (1 = shadowed) per light and pixel from maps[l] for light
 * l (count = scene->num_lights); a light whose map has no depth is fully
 * shadowed, as the reference's shadowMap == -1 case. */
luzrt_status luzrt_shadow_map_visibility(luzrt_ctx* ctx, const luzrt_gbuffer* g, const luzrt_scene* scene,
                                         const luzrt_shadow_map* maps, uint32_t count, uint8_t* shadow_out);
/* The light pass with shadow-map shadows and ray-traced AO (light.frag:main
 * with shadowType == Map). */
luzrt_status luzrt_light_pass_shadow_maps(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* g,
                                          const luzrt_scene* scene, const luzrt_shadow_map* maps,
                                          uint32_t count, float* rgba32f_out);

/* ------------------------------------------------------------ after the light pass */
/* Temporal accumulation of the light buffer (taa.comp:main, :271-307) and the
 * compose/tonemap of the result (present.frag:87-90, TonemapACES :27-35).
 * Matrices are column-major 4x4 (glm): view_proj is the current (jittered)
 * projection x view, prev_view_proj the previous frame's, inv_view_proj the
 * inverse of view_proj (used for background texels, whose depth is the far
 * plane).  Depth is the NDC depth of the position plane; images use the
 * reference's REPEAT addressing.  Set reconstruct to use the 3x3 Mitchell
 * reconstruction of the source (PostProcessingConstants.reconstruct). */
typedef struct luzrt_taa_desc {
    uint32_t width, height;
    const float* light;         /* device RGBA32F, this frame's light buffer */
    const float* history;       /* device RGBA32F, the previous TAA output */
    float* out;                 /* device RGBA32F (must not alias light or history) */
    const float* position;      /* device G-buffer position plane (float3) */
    const float* normal;        /* device G-buffer normal plane (zero = background) */
    float view_proj[16];
    float prev_view_proj[16];
    float inv_view_proj[16];
    float jitter[2];            /* NDC jitter of view_proj (CameraNode::GetJitter) */
    float prev_jitter[2];
    int32_t reconstruct;
    int32_t pad;
} luzrt_taa_desc;

luzrt_status luzrt_taa(luzrt_ctx* ctx, const luzrt_taa_desc* desc);
/* ACES tonemap + 1/2.2 gamma of the light buffer: RGBA8 UNORM (round half
 * up, alpha 255) and/or the FP32 tonemapped colour (alpha 1); either output
 * may be NULL. */
luzrt_status luzrt_compose(luzrt_ctx* ctx, uint32_t width, uint32_t height, const float* light_rgba32f,
                           uint32_t* rgba8_out, float* rgba32f_out);

/* ------------------------------------------------------------ batched frames */
/* Many independent scenes of uneven size and resolution rendered as one
 * frame (no counterpart in the reference, which renders one scene per
 * frame; each scene's result equals luzrt_primary_gbuffer + luzrt_light_pass
 * of that scene bit for bit).  Per frame: the BLASes registered with
 * luzrt_batch_add_refit are refit, every scene's TLAS is rebuilt from the
 * instances staged by luzrt_batch_set_instances (one launch for all TLASes),
 * G-buffers of scenes with a camera are rendered by primary rays, the shadow
 * and AO rays of ALL scenes are traced by one persistent traversal launch,
 * and every scene is shaded.  With use_graph != 0 the frame is captured once
 * into a hipGraph and replayed; it is recaptured when staged instance counts
 * or the trace configuration change.  Plain precision only.  The TLASes,
 * BLASes and buffers named here belong to the batch's frames until
 * luzrt_batch_destroy; the batch does not take ownership of them.  Two
 * scenes may not share a TLAS, an output buffer, or a G-buffer plane that a
 * camera writes (LUZRT_E_INVALID_ARG); G-buffers are tightly packed planes
 * with a position plane (LUZRT_E_UNSUPPORTED otherwise). */
typedef struct luzrt_batch luzrt_batch;

typedef struct luzrt_batch_scene {
    luzrt_tlas* tlas;                  /* rebuilt every frame from the staged instances */
    luzrt_gbuffer gbuffer;             /* device planes (written each frame when camera != NULL) */
    luzrt_scene scene;                 /* scene.frame + the frame index seeds the samples */
    const luzrt_camera* camera;        /* NULL: the caller supplies the G-buffer */
    const luzrt_material* materials;   /* host table indexed by instance custom_index */
    uint32_t material_count;
    uint8_t* shadow_counts;            /* device, [num_lights][width*height] */
    uint8_t* ao_counts;                /* device, [width*height] */
    float* rgba32f_out;                /* device, width*height*4 */
} luzrt_batch_scene;

luzrt_status luzrt_batch_create(luzrt_ctx* ctx, const luzrt_batch_scene* scenes, uint32_t count,
                                luzrt_batch** out);
/* Refit `blas` from `vertices` (device, the build's vertex stride) in every frame. */
luzrt_status luzrt_batch_add_refit(luzrt_ctx* ctx, luzrt_batch* batch, luzrt_blas* blas,
                                   const void* vertices);
/* Stage the instances of scene `scene` for the next frame (host array; waits
 * for the previous frame, whose staging copy reads the same memory). */
luzrt_status luzrt_batch_set_instances(luzrt_ctx* ctx, luzrt_batch* batch, uint32_t scene,
                                       const luzrt_instance* instances, uint32_t count);
/* Render one frame; asynchronous like the other calls. */
luzrt_status luzrt_batch_frame(luzrt_ctx* ctx, luzrt_batch* batch, uint32_t frame, int use_graph);
/* Waits for the last frame; its total ray count and any device error. */
luzrt_status luzrt_batch_rays(luzrt_ctx* ctx, luzrt_batch* batch, uint64_t* total_rays);
luzrt_status luzrt_batch_destroy(luzrt_ctx* ctx, luzrt_batch* batch);

/* ------------------------------------------------------------ learned denoiser */
/* A per-pixel MLP that denoises low-sample AO and per-light soft-shadow
 * visibility (no counterpart in the reference, whose light.frag uses the
 * noisy ray-query estimates directly).  One (pixel, light) sample has 150
 * features: a 5x5 neighbourhood (clamped to the image) x {AO fraction,
 * visibility of the light, relative depth clamp(16 (d_q - d_p) / d_p, -1, 1),
 * normal.xyz}, background neighbours (zero normal) giving {0, 0, 1, 0, 0, 0};
 * then 4 hidden layers of 64 (ReLU) and 2 outputs {AO, visibility}.  Weights
 * BF16 (MFMA), FP32 accumulation and biases; FP32 master copy for training.
 * Parameter order (FP32): W1[64][150] b1[64] W2[64][64] b2 W3 b3 W4 b4
 * W5[2][64] b5[2]. */
#define LUZRT_DN_FEATURES 150
#define LUZRT_DN_HIDDEN 64
#define LUZRT_DN_PARAMS 22274

typedef struct luzrt_denoiser luzrt_denoiser;

typedef struct luzrt_dn_frame {
    uint32_t width, height;
    const float* position;          /* device G-buffer planes (float3 per pixel) */
    const float* normal;
    float cam_pos[3];
    uint32_t num_lights;
    const uint8_t* shadow_counts;   /* device [num_lights][npix]: occluded samples */
    uint32_t shadow_samples;        /* samples per light behind shadow_counts */
    const uint8_t* ao_counts;       /* device [npix]: unoccluded AO samples */
    uint32_t ao_samples;
} luzrt_dn_frame;

/* batch_capacity: training samples per gradient (a multiple of 64, may be 0
 * for inference only).  Initial weights are He-uniform from `seed`. */
luzrt_status luzrt_denoiser_create(luzrt_ctx* ctx, uint32_t batch_capacity, uint64_t seed,
                                   luzrt_denoiser** out);
luzrt_status luzrt_denoiser_destroy(luzrt_ctx* ctx, luzrt_denoiser* dn);
luzrt_status luzrt_denoiser_get_params(luzrt_ctx* ctx, luzrt_denoiser* dn, float* host, uint32_t count);
/* Replaces the weights (LUZRT_DN_PARAMS floats) and resets the Adam state. */
luzrt_status luzrt_denoiser_set_params(luzrt_ctx* ctx, luzrt_denoiser* dn, const float* host,
                                       uint32_t count);
/* Denoised AO fraction [npix] and per-light visibility [num_lights][npix]
 * (device outputs in [0, 1]; 0 on background pixels). */
luzrt_status luzrt_denoise(luzrt_ctx* ctx, luzrt_denoiser* dn, const luzrt_dn_frame* frame,
                           float* ao_out, float* visibility_out);
/* The BF16 input features (raw bits, [count][150]) of explicit samples
 * (device array, pixel | light << 24), for inspection and external trainers. */
luzrt_status luzrt_denoiser_features(luzrt_ctx* ctx, const luzrt_dn_frame* frame, const uint32_t* samples,
                                     uint32_t count, uint16_t* features_out);
/* Training: forward and backward of `count` samples (device array, pixel |
 * light << 24; count a multiple of 64) of this frame against the targets
 * (device AO fraction [npix], visibility [num_lights][npix]).  The gradient
 * of the loss grad_scale/2 * sum (y - target)^2 is accumulated until
 * luzrt_denoiser_gradient. */
luzrt_status luzrt_denoiser_accumulate(luzrt_ctx* ctx, luzrt_denoiser* dn, const luzrt_dn_frame* frame,
                                       const float* target_ao, const float* target_visibility,
                                       const uint32_t* samples, uint32_t count, float grad_scale);
/* Weight gradient of the samples accumulated so far (deterministic order);
 * optional host copies of the gradient (LUZRT_DN_PARAMS floats) and of the
 * sum of squared errors (synchronise when requested).  Clears the batch. */
luzrt_status luzrt_denoiser_gradient(luzrt_ctx* ctx, luzrt_denoiser* dn, float* host_grad,
                                     float* loss_sum);
/* One Adam step with the last gradient; re-packs the BF16 weights. */
luzrt_status luzrt_denoiser_adam(luzrt_ctx* ctx, luzrt_denoiser* dn, float lr, float beta1, float beta2,
                                 float eps);

#ifdef __cplusplus
}
#endif
#endif /* LUZRT_H */
