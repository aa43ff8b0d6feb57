// shadowmap.hip — shadow maps on gfx950: the alternative visibility source of
// EvaluateShadow (light.frag:147-165) for shadowType == Map.
//
// The reference rasterises the maps (shadowMap.vert/.geom/.frag,
// DeferredRenderer.cpp:268-292).  Here a map texel holds the nearest surface
// through its centre, found by the closest-hit traversal the primary rays
// use: for a directional/spot light, the orthographic ray from the near
// (NDC z 0) to the far (NDC z 1) plane through the texel centre, whose hit
// parameter t is the NDC depth; for a point light, the ray from the light
// through the texel centre of a cube face, storing distance / z_far.  Texels
// without a hit keep the cleared depth 1.  The lookups follow light.frag:
// the light-space position of the shadow origin (no perspective divide,
// orthographic w = 1) or the light-to-fragment vector selects the texel,
// the depth is sampled bilinearly, and the comparisons are
// fragInLight.z >= depth and |lightToFrag| - 0.05 >= depth * z_far.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "lightpath.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

struct V4 { float x, y, z, w; };

// Column-major 4x4 times (x, y, z, w), summed left to right.
__device__ inline V4 mat4_vec(const float* m, float x, float y, float z, float w) {
    return V4{m[0] * x + m[4] * y + m[8] * z + m[12] * w, m[1] * x + m[5] * y + m[9] * z + m[13] * w,
              m[2] * x + m[6] * y + m[10] * z + m[14] * w, m[3] * x + m[7] * y + m[11] * z + m[15] * w};
}

// Direction through texel (i, j) of cube face f (Vulkan cube-map face
// selection inverted: +X -X +Y -Y +Z -Z).
__device__ inline V3 cube_dir(uint32_t f, uint32_t i, uint32_t j, uint32_t res) {
    float sc = ((float)i + 0.5f) / (float)res * 2.0f - 1.0f;
    float tc = ((float)j + 0.5f) / (float)res * 2.0f - 1.0f;
    switch (f) {
        case 0: return v3(1.0f, -tc, -sc);
        case 1: return v3(-1.0f, -tc, sc);
        case 2: return v3(sc, 1.0f, tc);
        case 3: return v3(sc, -1.0f, -tc);
        case 4: return v3(sc, -tc, 1.0f);
        default: return v3(-sc, -tc, -1.0f);
    }
}

// One ray per texel, persistent over the texels of every face.
__global__ __launch_bounds__(kBlock) void k_shadow_map(const Node* tlas_nodes, const InstRec* recs,
                                                       const InstAux* aux, luzrt_shadow_map m,
                                                       uint32_t* spill, uint32_t* error_flag, bool empty) {
    __shared__ uint32_t lds[kShort * kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = gridDim.x * kBlock;
    Stack st;
    st.init(lds + tid, kBlock, spill + (size_t)blockIdx.x * kBlock + tid, total);
    const uint32_t res = m.resolution;
    const bool point = m.light_type == LUZRT_LIGHT_POINT;
    const uint32_t n = (point ? 6u : 1u) * res * res;
    for (uint32_t k = blockIdx.x * kBlock + tid; k < n; k += total) {
        const uint32_t f = k / (res * res), r = k % (res * res), i = r % res, j = r / res;
        V3 o, d;
        float tmax;
        if (point) {
            o = v3(m.light_pos[0], m.light_pos[1], m.light_pos[2]);
            d = cube_dir(f, i, j, res);
            tmax = m.z_far;
        } else {
            float x = ((float)i + 0.5f) / (float)res * 2.0f - 1.0f;
            float y = ((float)j + 0.5f) / (float)res * 2.0f - 1.0f;
            V4 a = mat4_vec(m.inv_view_proj, x, y, 0.0f, 1.0f);
            V4 b = mat4_vec(m.inv_view_proj, x, y, 1.0f, 1.0f);
            o = v3(a.x, a.y, a.z);
            d = v3(b.x - a.x, b.y - a.y, b.z - a.z);
            tmax = 1.0f;
        }
        Best best;
        best.found = false;
        best.t = tmax;
        if (!empty) {
            RayState rs;
            ray_begin(rs, st, o, d, 0.0f, tmax);
            for (;;) {
                int res_step = ray_step<true>(rs, st, tlas_nodes, recs, aux, &best);
                if (res_step == kStepDone || (res_step == kStepContinue && rs.next == kNone && st.sp == 0))
                    break;
                if (res_step == kStepOverflow) { atomicOr(error_flag, 1u); break; }
            }
        }
        float depth = 1.0f;
        if (best.found) {
            if (point) {
                float dist = best.t * vlen(d);
                float v = dist / m.z_far;
                depth = v < 1.0f ? v : 1.0f;
            } else {
                depth = best.t;
            }
        }
        m.depth[k] = depth;
    }
}

__device__ inline int wrapi(int v, int n) { int r = v % n; return r < 0 ? r + n : r; }
__device__ inline int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Bilinear sample at normalised (u, v) of a res x res depth image: REPEAT
// addressing (2D maps) or clamp (cube faces).
__device__ inline float sample_depth(const float* img, int res, float u, float v, bool repeat) {
    float tx = u * (float)res - 0.5f, ty = v * (float)res - 0.5f;
    float fx = floorf(tx), fy = floorf(ty);
    float ax = tx - fx, ay = ty - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = repeat ? wrapi(x0, res) : clampi(x0, res), xb = repeat ? wrapi(x0 + 1, res) : clampi(x0 + 1, res);
    int ya = repeat ? wrapi(y0, res) : clampi(y0, res), yb = repeat ? wrapi(y0 + 1, res) : clampi(y0 + 1, res);
    float d00 = img[(size_t)ya * res + xa], d10 = img[(size_t)ya * res + xb];
    float d01 = img[(size_t)yb * res + xa], d11 = img[(size_t)yb * res + xb];
    float top = d00 * (1.0f - ax) + d10 * ax;
    float bot = d01 * (1.0f - ax) + d11 * ax;
    return top * (1.0f - ay) + bot * ay;
}

// EvaluateShadow, shadowType == Map: 1 when shadowed.
__device__ inline uint8_t map_shadow(const luzrt_shadow_map& m, V3 P, V3 so) {
    if (!m.depth) return 1;
    const int res = (int)m.resolution;
    if (m.light_type == LUZRT_LIGHT_POINT) {
        V3 l = v3(P.x - m.light_pos[0], P.y - m.light_pos[1], P.z - m.light_pos[2]);
        float ax = fabsf(l.x), ay = fabsf(l.y), az = fabsf(l.z);
        uint32_t f;
        float sc, tc, ma;
        if (ax >= ay && ax >= az) { f = l.x >= 0.0f ? 0u : 1u; ma = ax; sc = l.x >= 0.0f ? -l.z : l.z; tc = -l.y; }
        else if (ay >= az) { f = l.y >= 0.0f ? 2u : 3u; ma = ay; sc = l.x; tc = l.y >= 0.0f ? l.z : -l.z; }
        else { f = l.z >= 0.0f ? 4u : 5u; ma = az; sc = l.z >= 0.0f ? l.x : -l.x; tc = -l.y; }
        float u = (sc / ma + 1.0f) * 0.5f, v = (tc / ma + 1.0f) * 0.5f;
        float depth = sample_depth(m.depth + (size_t)f * res * res, res, u, v, false);
        return vlen(l) - 0.05f >= depth * m.z_far ? 1 : 0;
    }
    V4 q = mat4_vec(m.view_proj, so.x, so.y, so.z, 1.0f);
    float depth = sample_depth(m.depth, res, q.x * 0.5f + 0.5f, q.y * 0.5f + 0.5f, true);
    return q.z >= depth ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_map_visibility(luzrt_gbuffer g, const DevScene* ds,
                                                           const luzrt_shadow_map* maps, uint32_t nmaps,
                                                           uint8_t* shadow) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t npix = g.width * g.height;
    if (p >= npix) return;
    const luzrt_scene& S = ds->s;
    const uint32_t px = p % g.width, py = p / g.width;
    V3 N = gb_normal(g, px, py);
    const bool fg = vlen(N) != 0.0f;
    V3 P = gb_position(g, px, py);
    // shadow origin (light.frag:138-139)
    V3 cam = v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2]);
    float bias = fmax_(vlen(vsub(P, cam)) * 0.01f, 0.05f);
    V3 so = vadd(P, vmul(N, bias));
    for (int l = 0; l < S.num_lights; ++l) {
        uint8_t sh = 0;
        if (fg) sh = (uint32_t)l < nmaps ? map_shadow(maps[l], P, so) : 1;
        shadow[(size_t)l * npix + p] = sh;
    }
}

}  // namespace
}  // namespace luz

using namespace luz;

extern "C" luzrt_status luzrt_shadow_map_build(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_shadow_map* m) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !m || !m->depth) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shadow_map_build: null argument");
    if (m->resolution == 0 || m->resolution > 16384)
        return fail(ctx, LUZRT_E_LIMIT, "shadow map resolution %u outside 1..16384", m->resolution);
    if (m->light_type < 0 || m->light_type > 2) return fail(ctx, LUZRT_E_INVALID_ARG, "bad light type");
    if (m->light_type == LUZRT_LIGHT_POINT && !(m->z_far > 0.0f))
        return fail(ctx, LUZRT_E_INVALID_ARG, "point-light shadow map needs z_far > 0");
    luzrt_status s = ensure(ctx, ctx->counters, kCounterBytes);
    if (s != LUZRT_OK) return s;
    if ((s = tlas_acquire(ctx, tlas)) != LUZRT_OK) return s;
    uint32_t groups = (uint32_t)ctx->num_cus * 4;
    if (groups == 0) groups = 1;
    if ((s = ensure(ctx, ctx->spill, (size_t)groups * kBlock * kSpill * 4)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    stage_begin(ctx, kStagePrimary);
    k_shadow_map<<<groups, kBlock, 0, ctx->stream>>>(tlas->nodes, tlas->recs, tlas->aux, *m,
                                                    (uint32_t*)ctx->spill.ptr, counters + kCntError,
                                                    tlas->count == 0);
    stage_end(ctx, kStagePrimary);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

extern "C" luzrt_status luzrt_shadow_map_visibility(luzrt_ctx* ctx, const luzrt_gbuffer* g, const luzrt_scene* scene,
                                                    const luzrt_shadow_map* maps, uint32_t count, uint8_t* shadow) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!g || !scene || (count && !maps) || (scene->num_lights > 0 && !shadow))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shadow_map_visibility: null argument");
    DevScene ds;
    luzrt_status s = upload_scene(ctx, scene, &ds);
    if (s != LUZRT_OK) return s;
    luzrt_gbuffer gl;
    if ((s = check_gbuffer(ctx, g, &gl)) != LUZRT_OK) return s;
    if (scene->num_lights == 0) return LUZRT_OK;
    for (uint32_t l = 0; l < count; ++l)
        if (maps[l].depth && (maps[l].resolution == 0 || maps[l].light_type < 0 || maps[l].light_type > 2))
            return fail(ctx, LUZRT_E_INVALID_ARG, "shadow map %u is malformed", l);
    // the map descriptors travel to the device for the lookup kernel
    luzrt_shadow_map* dmaps = nullptr;
    const uint32_t n = count ? count : 1;
    LUZ_HIP(ctx, hipMallocAsync((void**)&dmaps, n * sizeof(luzrt_shadow_map), ctx->stream));
    if (count)
        LUZ_HIP(ctx, hipMemcpyAsync(dmaps, maps, count * sizeof(luzrt_shadow_map), hipMemcpyHostToDevice, ctx->stream));
    const uint32_t npix = g->width * g->height;
    k_map_visibility<<<(npix + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(gl, (const DevScene*)ctx->scene.ptr,
                                                                              dmaps, count, shadow);
    LUZ_HIP(ctx, hipGetLastError());
    LUZ_HIP(ctx, hipFreeAsync(dmaps, ctx->stream));
    // the host array must stay valid until the copy has run
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

extern "C" luzrt_status luzrt_light_pass_shadow_maps(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* g,
                                                     const luzrt_scene* scene, const luzrt_shadow_map* maps,
                                                     uint32_t count, float* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !g || !scene || !out) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_light_pass_shadow_maps: null argument");
    luzrt_gbuffer gl;
    luzrt_status s = check_gbuffer(ctx, g, &gl);
    if (s != LUZRT_OK) return s;
    const size_t npix = (size_t)g->width * g->height;
    const int nl = scene->num_lights < 0 ? 0 : scene->num_lights;
    if ((s = ensure(ctx, ctx->vis_shadow, npix * (size_t)(nl > 0 ? nl : 1))) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->vis_ao, npix)) != LUZRT_OK) return s;
    // AO rays only: every light's shadow comes from its map (1 "sample" each)
    luzrt_scene ao_only = *scene;
    for (int l = 0; l < nl; ++l) ao_only.lights[l].num_shadow_samples = 0;
    s = luzrt_trace_visibility(ctx, tlas, g, &ao_only, (uint8_t*)ctx->vis_shadow.ptr, (uint8_t*)ctx->vis_ao.ptr);
    if (s != LUZRT_OK) return s;
    s = luzrt_shadow_map_visibility(ctx, g, scene, maps, count, (uint8_t*)ctx->vis_shadow.ptr);
    if (s != LUZRT_OK) return s;
    luzrt_scene shade = *scene;
    for (int l = 0; l < nl; ++l) shade.lights[l].num_shadow_samples = 1;
    return luzrt_shade(ctx, g, &shade, (const uint8_t*)ctx->vis_shadow.ptr, (const uint8_t*)ctx->vis_ao.ptr, out);
}
