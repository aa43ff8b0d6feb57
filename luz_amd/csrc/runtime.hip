// runtime.hip — context, memory, acceleration-structure handles and errors of
// the luzrt C ABI.
//
// Replaces the parts of the Vulkan wrapper the path needs: the device/queue
// context (VulkanWrapper.cpp:243 `static Context _ctx`), buffer creation and
// staging copies (:422-486), the BLAS/TLAS handle objects (:657-836) and the
// GPU timestamps (:1276-1332).  Unlike the reference, failures are returned
// as status codes with a message instead of ASSERT/abort (Base.hpp:51-59).
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <algorithm>

#include "internal.hpp"
#include "traverse.hpp"
#include "cull.hpp"

namespace luz {

luzrt_status fail(luzrt_ctx* ctx, luzrt_status s, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->error = buf;
    }
    return s;
}

luzrt_status hip_fail(luzrt_ctx* ctx, hipError_t e, const char* what) {
    luzrt_status s = e == hipErrorOutOfMemory ? LUZRT_E_OUT_OF_MEMORY : LUZRT_E_DEVICE;
    return fail(ctx, s, "%s failed: %s", what, hipGetErrorString(e));
}

luzrt_status ensure(luzrt_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.ptr) return LUZRT_OK;
    if (b.ptr) {
        LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LUZ_HIP(ctx, hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t sz = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.ptr, sz);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return hip_fail(ctx, e, "hipMalloc(workspace)");
    }
    b.bytes = sz;
    return LUZRT_OK;
}

namespace {
__global__ void k_zero_fill(uint8_t* p, size_t bytes) {
    const size_t words = bytes / 4;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint32_t*>(p)[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < (bytes & 3)) p[words * 4 + threadIdx.x] = 0;
}
}  // namespace

luzrt_status zero_fill(luzrt_ctx* ctx, void* p, size_t bytes) {
    if (bytes == 0) return LUZRT_OK;
    if (((uintptr_t)p & 3) != 0) return fail(ctx, LUZRT_E_INVALID_ARG, "zero_fill: unaligned pointer");
    size_t words = bytes / 4;
    uint32_t blocks = (uint32_t)((words + 255) / 256);
    if (blocks == 0) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    k_zero_fill<<<blocks, 256, 0, ctx->stream>>>((uint8_t*)p, bytes);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

void stage_begin(luzrt_ctx* ctx, Stage s, hipStream_t st) {
    if (!ctx->timing) return;
    hipEventRecord(ctx->ev[s][0], st ? st : ctx->stream);
    ctx->ev_used[s] = true;
}
void stage_end(luzrt_ctx* ctx, Stage s, hipStream_t st) {
    if (ctx->timing) hipEventRecord(ctx->ev[s][1], st ? st : ctx->stream);
}

luzrt_status tlas_acquire(luzrt_ctx* ctx, const luzrt_tlas* tlas) {
    luzrt_tlas* t = const_cast<luzrt_tlas*>(tlas);
    ctx->release_clean = false;   // this call reads a TLAS after any earlier release
    if (!t->build_pending) return LUZRT_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LUZ_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        LUZ_HIP(ctx, hipEventSynchronize(t->built));
    else
        LUZ_HIP(ctx, hipStreamWaitEvent(ctx->stream, t->built, 0));
    t->build_pending = false;
    ctx->pending_builds--;
    return LUZRT_OK;
}

luzrt_status tlas_release(luzrt_ctx* ctx) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LUZ_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return LUZRT_OK;
    LUZ_HIP(ctx, hipEventRecord(ctx->tlas_release, ctx->stream));
    ctx->release_clean = true;
    return LUZRT_OK;
}

luzrt_status build_stream_sync(luzrt_ctx* ctx) {
    if (ctx->build_stream) LUZ_HIP(ctx, hipStreamSynchronize(ctx->build_stream));
    return LUZRT_OK;
}

luzrt_status check_gbuffer(luzrt_ctx* ctx, const luzrt_gbuffer* g, luzrt_gbuffer* out) {
    if (g->width == 0 || g->height == 0 || g->width > 16384 || g->height > 16384)
        return fail(ctx, LUZRT_E_LIMIT, "G-buffer size %ux%u outside 1..16384", g->width, g->height);
    if ((!g->position && !g->depth) || !g->normal || !g->albedo || !g->material || !g->emission)
        return fail(ctx, LUZRT_E_INVALID_ARG, "G-buffer plane is null");
    luzrt_gbuffer L = *g;
    const uint64_t W = g->width;
    auto stride = [&](uint32_t& s, const char* name) {
        if (s == 0) s = 12;
        if (s != 12 && s != 16)
            return fail(ctx, LUZRT_E_INVALID_ARG, "G-buffer %s stride %u is neither 12 nor 16", name, s);
        return LUZRT_OK;
    };
    auto pitch = [&](uint32_t& p, uint64_t row, const void* base, const char* name) {
        if (p == 0) p = (uint32_t)row;
        if (p < row || (p & 3u) != 0 || ((uintptr_t)base & 3u) != 0)
            return fail(ctx, LUZRT_E_INVALID_ARG,
                        "G-buffer %s: pitch %u below the row (%llu bytes) or not 4-byte aligned", name, p,
                        (unsigned long long)row);
        return LUZRT_OK;
    };
    luzrt_status s;
    if ((s = stride(L.position_stride, "position")) != LUZRT_OK) return s;
    if ((s = stride(L.normal_stride, "normal")) != LUZRT_OK) return s;
    if (W * 16 > 0xFFFFFFFFull) return fail(ctx, LUZRT_E_LIMIT, "G-buffer row too wide");
    if (L.position) {
        if ((s = pitch(L.position_pitch, W * L.position_stride, L.position, "position")) != LUZRT_OK) return s;
        L.depth = nullptr;
    } else if ((s = pitch(L.depth_pitch, W * 4, L.depth, "depth")) != LUZRT_OK) {
        return s;
    }
    if ((s = pitch(L.normal_pitch, W * L.normal_stride, L.normal, "normal")) != LUZRT_OK) return s;
    if ((s = pitch(L.albedo_pitch, W * 4, L.albedo, "albedo")) != LUZRT_OK) return s;
    if ((s = pitch(L.material_pitch, W * 4, L.material, "material")) != LUZRT_OK) return s;
    if ((s = pitch(L.emission_pitch, W * 4, L.emission, "emission")) != LUZRT_OK) return s;
    *out = L;
    return LUZRT_OK;
}

bool gbuffer_packed(const luzrt_gbuffer& g) {
    const uint32_t W = g.width;
    return g.position && g.position_stride == 12 && g.normal_stride == 12 && g.position_pitch == W * 12 &&
           g.normal_pitch == W * 12 && g.albedo_pitch == W * 4 && g.material_pitch == W * 4 &&
           g.emission_pitch == W * 4;
}

luzrt_status fill_scene(luzrt_ctx* ctx, const luzrt_scene* scene, DevScene* ds, bool trace) {
    if (scene->num_lights < 0 || scene->num_lights > LUZRT_MAX_LIGHTS)
        return fail(ctx, LUZRT_E_LIMIT, "num_lights %d outside 0..%d", scene->num_lights,
                    LUZRT_MAX_LIGHTS);
    if (scene->ao_num_samples < 0 || scene->ao_num_samples > LUZRT_MAX_SAMPLES)
        return fail(ctx, LUZRT_E_LIMIT, "ao_num_samples %d outside 0..%d", scene->ao_num_samples,
                    LUZRT_MAX_SAMPLES);
    if (scene->shadow_type < LUZRT_SHADOW_DISABLED || scene->shadow_type > LUZRT_SHADOW_MAP)
        return fail(ctx, LUZRT_E_INVALID_ARG, "shadow_type %d is not one of 0, 1, 2", scene->shadow_type);
    std::memset(ds, 0, sizeof(*ds));
    ds->s = *scene;
    uint32_t j = 0;
    for (int l = 0; l < scene->num_lights; ++l) {
        const luzrt_light& L = scene->lights[l];
        if (L.type < 0 || L.type > 2)
            return fail(ctx, LUZRT_E_INVALID_ARG, "light %d has unknown type %d", l, L.type);
        if (L.num_shadow_samples < 0 || L.num_shadow_samples > LUZRT_MAX_SAMPLES)
            return fail(ctx, LUZRT_E_LIMIT, "light %d num_shadow_samples %d outside 0..%d", l,
                        L.num_shadow_samples, LUZRT_MAX_SAMPLES);
        // shadow rays only in ray-traced scenes (GPUScene.cpp:248 packs
        // numShadowSamples = 0 otherwise; light.frag:141)
        if (trace && scene->shadow_type != LUZRT_SHADOW_RAYTRACED) ds->s.lights[l].num_shadow_samples = 0;
        ds->slot_base[l] = j;
        j += (uint32_t)ds->s.lights[l].num_shadow_samples;
    }
    ds->r_shadow = j;
    ds->r_total = j + (uint32_t)scene->ao_num_samples;
    return LUZRT_OK;
}

namespace {
// The live part of a scene block (its first num_lights lights and everything
// after the light array) as kernel arguments: words [0, n0) and
// [off1, off1 + n1) of the block.
constexpr uint32_t kScenePutWords = 480;
struct ScenePut {
    uint32_t n0, off1, n1, pad;
    uint32_t w[kScenePutWords];
};
__global__ void k_put_scene_live(uint32_t* dst, ScenePut p) {
    for (uint32_t i = threadIdx.x; i < p.n0 + p.n1; i += blockDim.x)
        dst[i < p.n0 ? i : p.off1 + (i - p.n0)] = p.w[i];
}
}  // namespace

luzrt_status upload_scene(luzrt_ctx* ctx, const luzrt_scene* scene, DevScene* ds, uint32_t ray_group,
                          bool trace) {
    luzrt_status s = fill_scene(ctx, scene, ds, trace);
    if (s != LUZRT_OK) return s;
    ds->ray_group = ray_group;
    s = ensure(ctx, ctx->scene, sizeof(DevScene));
    if (s != LUZRT_OK) return s;
    // Kernels read lights 0 .. num_lights - 1 only.  When the live part fits
    // the kernel's arguments it goes in a kernel launch, which the next kernel
    // follows without the gap (about 6 us) a staged copy leaves before it.
    static_assert(offsetof(DevScene, s) == 0 && offsetof(luzrt_scene, lights) == 0, "light array first");
    static_assert(sizeof(DevScene) % 4 == 0 && sizeof(luzrt_light) % 4 == 0, "whole words");
    const uint32_t n0 = (uint32_t)(ds->s.num_lights * sizeof(luzrt_light) / 4);
    const uint32_t off1 = (uint32_t)(LUZRT_MAX_LIGHTS * sizeof(luzrt_light) / 4);
    const uint32_t n1 = (uint32_t)(sizeof(DevScene) / 4) - off1;
    if (n0 + n1 > kScenePutWords) return stage_upload(ctx, ctx->scene.ptr, ds, sizeof(DevScene));
    ScenePut p;
    std::memset(&p, 0, sizeof(p));
    p.n0 = n0;
    p.off1 = off1;
    p.n1 = n1;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(ds);
    std::memcpy(p.w, src, (size_t)n0 * 4);
    std::memcpy(p.w + n0, src + off1, (size_t)n1 * 4);
    k_put_scene_live<<<1, 256, 0, ctx->stream>>>(reinterpret_cast<uint32_t*>(ctx->scene.ptr), p);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

luzrt_status stage_upload(luzrt_ctx* ctx, void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return LUZRT_OK;
    if (!st) st = ctx->stream;
    luzrt_ctx::Staging& sl = ctx->staging[ctx->staging_next];
    ctx->staging_next = (ctx->staging_next + 1) % luzrt_ctx::kStagingSlots;
    if (sl.pending) {
        LUZ_HIP(ctx, hipEventSynchronize(sl.done));
        sl.pending = false;
    }
    if (!sl.done) LUZ_HIP(ctx, hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    if (sl.bytes < bytes) {
        if (sl.host) LUZ_HIP(ctx, hipHostFree(sl.host));
        sl.host = nullptr;
        sl.bytes = 0;
        const size_t sz = bytes < 4096 ? 4096 : bytes;
        LUZ_HIP(ctx, hipHostMalloc(&sl.host, sz, hipHostMallocDefault));
        sl.bytes = sz;
    }
    std::memcpy(sl.host, src, bytes);
    LUZ_HIP(ctx, hipMemcpyAsync(dst, sl.host, bytes, hipMemcpyHostToDevice, st));
    LUZ_HIP(ctx, hipEventRecord(sl.done, st));
    sl.pending = true;
    return LUZRT_OK;
}

// Waits for the build and reports (then clears) the device build-error flag:
// 2 = a vertex index >= vertex_count, 1 = malformed hierarchy.
static luzrt_status check_build_flag(luzrt_ctx* ctx) {
    uint32_t flag = 0;
    uint32_t* d = (uint32_t*)ctx->counters.ptr + kCntBuildError;
    LUZ_HIP(ctx, hipMemcpyAsync(&flag, d, 4, hipMemcpyDeviceToHost, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!flag) return LUZRT_OK;
    LUZ_HIP(ctx, hipMemsetAsync(d, 0, 4, ctx->stream));
    if (flag & 2u) return fail(ctx, LUZRT_E_INVALID_ARG, "a vertex index is >= vertex_count");
    return fail(ctx, LUZRT_E_DEVICE, "BVH build produced a malformed hierarchy");
}

// FP64 inverse of a 3x4 affine transform, rounded once to FP32.
static bool invert_3x4(const float m[12], float out[12], float nmat[9]) {
    double r00 = m[0], r01 = m[1], r02 = m[2], t0 = m[3];
    double r10 = m[4], r11 = m[5], r12 = m[6], t1 = m[7];
    double r20 = m[8], r21 = m[9], r22 = m[10], t2 = m[11];
    double c00 = r11 * r22 - r12 * r21;
    double c01 = r12 * r20 - r10 * r22;
    double c02 = r10 * r21 - r11 * r20;
    double det = r00 * c00 + r01 * c01 + r02 * c02;
    if (det == 0.0 || !std::isfinite(det)) return false;
    double inv[3][3];
    inv[0][0] = c00 / det;
    inv[1][0] = c01 / det;
    inv[2][0] = c02 / det;
    inv[0][1] = (r02 * r21 - r01 * r22) / det;
    inv[1][1] = (r00 * r22 - r02 * r20) / det;
    inv[2][1] = (r01 * r20 - r00 * r21) / det;
    inv[0][2] = (r01 * r12 - r02 * r11) / det;
    inv[1][2] = (r02 * r10 - r00 * r12) / det;
    inv[2][2] = (r00 * r11 - r01 * r10) / det;
    double tv[3] = {t0, t1, t2};
    for (int i = 0; i < 3; ++i) {
        double ti = -(inv[i][0] * tv[0] + inv[i][1] * tv[1] + inv[i][2] * tv[2]);
        for (int j = 0; j < 3; ++j) out[i * 4 + j] = (float)inv[i][j];
        out[i * 4 + 3] = (float)ti;
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) nmat[r * 3 + c] = out[c * 4 + r];
    return true;
}

// Grouping a pixel's rays keeps its G-buffer texel in cache but puts every
// light's rays in flight at once; ordering by slot keeps the rays in flight
// to one light sample, which the BVH reads favour once the BVHs outgrow the
// caches.  Measured (trace kernel): 1080p sphere field, 11 MB of BVH, 9 lights
// + 16 AO rays: 1.48 ms grouped, 1.73 ms by slot; 4K terrain, 8-wide BVH of
// 4.2 M triangles, 36 lights + 8 AO rays: 32.8 ms grouped, 29.3 ms by slot.
constexpr size_t kGroupedBvhBytes = 32u << 20;   // the L2 of all XCDs together

uint32_t ray_group_for(const luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_scene* scene) {
    if (ctx->ray_order == LUZRT_RAY_ORDER_SLOT) return 0;
    if (ctx->ray_order == LUZRT_RAY_ORDER_GROUPED) return kRayGroup;
    // a call that can cull its ray groups (which needs the grouped order)
    // gains more from the culling than the slot order's coherent BVH reads
    // give: on the 4K terrain (a 270 MB BVH) 20.7 against 21.9 ms
    if (scene && ctx->group_cull && tlas->count > 0 && tlas->count <= kMaxCullInst &&
        scene->num_lights >= 0 && scene->num_lights <= LUZRT_MAX_LIGHTS) {
        uint32_t rays = scene->ao_num_samples > 0 ? (uint32_t)scene->ao_num_samples : 0u;
        if (scene->shadow_type == LUZRT_SHADOW_RAYTRACED)
            for (int l = 0; l < scene->num_lights; ++l)
                rays += scene->lights[l].num_shadow_samples > 0 ? (uint32_t)scene->lights[l].num_shadow_samples : 0u;
        if (rays > 0 && rays <= kMaxCullSlots) return kRayGroup;
    }
    std::vector<const luzrt_blas*> seen(tlas->referenced);
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    size_t bytes = 0;
    for (const luzrt_blas* b : seen)
        bytes += tlas->wide ? (size_t)b->wide_count * sizeof(Node8) + (size_t)b->tri_count * 36
                            : (size_t)b->node_count * sizeof(Node) + (size_t)b->tri_count * 36;
    return bytes <= kGroupedBvhBytes ? kRayGroup : 0;
}

}  // namespace luz

using namespace luz;

extern "C" {

const char* luzrt_status_string(luzrt_status s) {
    switch (s) {
        case LUZRT_OK: return "ok";
        case LUZRT_E_INVALID_ARG: return "invalid argument";
        case LUZRT_E_OUT_OF_MEMORY: return "out of device memory";
        case LUZRT_E_DEVICE: return "device error";
        case LUZRT_E_IN_USE: return "object in use";
        case LUZRT_E_LIMIT: return "limit exceeded";
        case LUZRT_E_NO_DEVICE: return "no HIP device";
        case LUZRT_E_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char* luzrt_last_error(const luzrt_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

luzrt_status luzrt_create(luzrt_ctx** out, int device) {
    if (!out) return LUZRT_E_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return LUZRT_E_NO_DEVICE;
    if (device < 0 || device >= n) return LUZRT_E_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return LUZRT_E_DEVICE;
    luzrt_ctx* ctx = new luzrt_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return LUZRT_E_DEVICE; }
    ctx->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return LUZRT_E_DEVICE;
    }
    if (hipStreamCreateWithFlags(&ctx->build_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->build_mark, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->tlas_release, hipEventDisableTiming) != hipSuccess) {
        luzrt_destroy(ctx);
        return LUZRT_E_DEVICE;
    }
    // Stage events only time: luzrt_get_timings reads them after a full stream
    // synchronisation, so they skip the system-scope fence (an L2 write-back and
    // invalidate between the stages' kernels, which cost about 10 us per record).
    for (int s = 0; s < kStageCount; ++s)
        for (int k = 0; k < 2; ++k) hipEventCreateWithFlags(&ctx->ev[s][k], hipEventDisableSystemFence);
    if (ensure(ctx, ctx->counters, kCounterBytes) != LUZRT_OK ||
        hipMemset(ctx->counters.ptr, 0, kCounterBytes) != hipSuccess) {
        luzrt_destroy(ctx);
        return LUZRT_E_DEVICE;
    }
    static_scratch_init(ctx);
    *out = ctx;
    return LUZRT_OK;
}

luzrt_status luzrt_destroy(luzrt_ctx* ctx) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (ctx->build_stream) hipStreamSynchronize(ctx->build_stream);
    DevBuf* bufs[] = {&ctx->hits, &ctx->active, &ctx->act_index, &ctx->act_rec,
                      &ctx->counters, &ctx->spill, &ctx->lut, &ctx->scene, &ctx->vis_shadow,
                      &ctx->vis_ao, &ctx->cull_cand, &ctx->cull_bits, &ctx->cull_rows};
    for (DevBuf* b : bufs)
        if (b->ptr) hipFree(b->ptr);
    for (int s = 0; s < kStageCount; ++s)
        for (int k = 0; k < 2; ++k) hipEventDestroy(ctx->ev[s][k]);
    for (luzrt_ctx::Staging& sl : ctx->staging) {
        if (sl.host) hipHostFree(sl.host);
        if (sl.done) hipEventDestroy(sl.done);
    }
    if (ctx->build_mark) hipEventDestroy(ctx->build_mark);
    if (ctx->tlas_release) hipEventDestroy(ctx->tlas_release);
    if (ctx->build_stream) hipStreamDestroy(ctx->build_stream);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return LUZRT_OK;
}

// After a stream synchronisation: reports (and clears) the traversal error
// latch, which a ray that overflows its stack or step limit sets on the
// device and no kernel clears.
static luzrt_status check_trace_latch(luzrt_ctx* ctx) {
    if (!ctx->counters.ptr) return LUZRT_OK;
    uint32_t* d = (uint32_t*)ctx->counters.ptr + kCntError;
    uint32_t flag = 0;
    LUZ_HIP(ctx, hipMemcpy(&flag, d, 4, hipMemcpyDeviceToHost));
    if (!flag) return LUZRT_OK;
    const uint32_t zero = 0;
    LUZ_HIP(ctx, hipMemcpy(d, &zero, 4, hipMemcpyHostToDevice));
    return fail(ctx, LUZRT_E_LIMIT,
                "a ray exceeded the traversal stack (%u LDS + %d spilled entries) or its step limit "
                "since the last check; such rays were written as misses", ctx->stack_entries, kSpill);
}

luzrt_status luzrt_sync(luzrt_ctx* ctx) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (build_stream_sync(ctx) != LUZRT_OK) return LUZRT_E_DEVICE;
    return check_trace_latch(ctx);
}

luzrt_status luzrt_get_timings(luzrt_ctx* ctx, luzrt_timings* out) {
    if (!ctx || !out) return LUZRT_E_INVALID_ARG;
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (build_stream_sync(ctx) != LUZRT_OK) return LUZRT_E_DEVICE;
    float* f[kStageCount] = {&out->build_ms, &out->compact_ms, &out->raygen_ms, &out->trace_ms,
                             &out->resolve_ms, &out->shade_ms, &out->primary_ms, &out->denoise_ms,
                             &out->post_ms};
    for (int s = 0; s < kStageCount; ++s) {
        *f[s] = 0.0f;
        if (ctx->ev_used[s]) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ctx->ev[s][0], ctx->ev[s][1]) == hipSuccess) *f[s] = ms;
        }
        ctx->ev_used[s] = false;
    }
    out->rays = 0;
    out->active_pixels = 0;
    out->traced_rays = 0;
    if (ctx->counters.ptr) {
        uint32_t c[kCntCount];
        LUZ_HIP(ctx, hipMemcpy(c, ctx->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
        out->rays = c[kCntRays];
        out->active_pixels = c[kCntActive];
        out->traced_rays = ctx->last_culled ? (uint64_t)c[kCntLive] * 64u : out->rays;
        if (c[kCntError]) return check_trace_latch(ctx);
        if (c[kCntBuildError])
            return fail(ctx, LUZRT_E_INVALID_ARG, "acceleration-structure build error flag %u",
                        c[kCntBuildError]);
    }
    return LUZRT_OK;
}

luzrt_status luzrt_set_trace_config(luzrt_ctx* ctx, uint32_t groups_per_cu, uint32_t refill,
                                    uint32_t stack_entries, uint32_t tile_width) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (groups_per_cu > 64 || refill > 64)
        return fail(ctx, LUZRT_E_INVALID_ARG, "trace config out of range (groups<=64, refill<=64)");
    if (stack_entries && stack_entries != 8 && stack_entries != 16 && stack_entries != 32)
        return fail(ctx, LUZRT_E_INVALID_ARG, "stack_entries must be 8, 16 or 32");
    if (tile_width && tile_width != 8 && tile_width != 16 && tile_width != 32)
        return fail(ctx, LUZRT_E_INVALID_ARG, "tile_width must be 8, 16 or 32");
    if (groups_per_cu) ctx->groups_per_cu = groups_per_cu;
    if (refill) ctx->refill_threshold = refill;
    if (stack_entries) ctx->stack_entries = stack_entries;
    if (tile_width) ctx->tile_width = tile_width;
    return LUZRT_OK;
}

luzrt_status luzrt_set_stage_timing(luzrt_ctx* ctx, int enable) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    ctx->timing = enable != 0;
    return LUZRT_OK;
}

luzrt_status luzrt_set_group_culling(luzrt_ctx* ctx, int enable) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    ctx->group_cull = enable != 0;
    return LUZRT_OK;
}

luzrt_status luzrt_device_info(const luzrt_ctx* ctx, uint32_t* compute_units, uint32_t* clock_khz) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return LUZRT_E_DEVICE;
    if (compute_units) *compute_units = (uint32_t)prop.multiProcessorCount;
    if (clock_khz) *clock_khz = (uint32_t)prop.clockRate;
    return LUZRT_OK;
}

luzrt_status luzrt_set_step_limit(luzrt_ctx* ctx, uint32_t max_steps) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    ctx->max_steps = max_steps ? max_steps : kMaxSteps;
    return LUZRT_OK;
}

luzrt_status luzrt_set_ray_order(luzrt_ctx* ctx, uint32_t order) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (order != LUZRT_RAY_ORDER_AUTO && order != LUZRT_RAY_ORDER_SLOT && order != LUZRT_RAY_ORDER_GROUPED)
        return fail(ctx, LUZRT_E_INVALID_ARG, "unknown ray order %u", order);
    ctx->ray_order = order;
    return LUZRT_OK;
}

luzrt_status luzrt_set_precision(luzrt_ctx* ctx, int mode) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (mode != LUZRT_PRECISION_PLAIN && mode != LUZRT_PRECISION_COMPACT && mode != LUZRT_PRECISION_WIDE)
        return fail(ctx, LUZRT_E_INVALID_ARG, "unknown precision mode %d", mode);
    ctx->precision = mode;
    return LUZRT_OK;
}

luzrt_status luzrt_blas_wide_arrays(luzrt_ctx* ctx, luzrt_blas* b, const void** nodes,
                                    uint32_t* node_count, const float** tris,
                                    const uint32_t** prims, uint32_t* tri_count) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_blas_wide_arrays: null BLAS");
    if (b->tri_count > 0 && !b->wide_valid) {
        luzrt_status s = blas_wide_device(ctx, b);
        if (s != LUZRT_OK) return s;
    }
    if (nodes) *nodes = b->wide;
    if (node_count) *node_count = b->tri_count > 0 ? b->wide_count : 0;
    if (tris) *tris = b->tris8;
    if (prims) *prims = b->prims8;
    if (tri_count) *tri_count = b->tri_count;
    return LUZRT_OK;
}

luzrt_status luzrt_malloc(luzrt_ctx* ctx, size_t bytes, void** p) {
    if (!ctx || !p) return LUZRT_E_INVALID_ARG;
    LUZ_HIP(ctx, hipMalloc(p, bytes ? bytes : 1));
    return LUZRT_OK;
}

luzrt_status luzrt_free(luzrt_ctx* ctx, void* p) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!p) return LUZRT_OK;
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LUZ_HIP(ctx, hipFree(p));
    return LUZRT_OK;
}

luzrt_status luzrt_upload(luzrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return LUZRT_E_INVALID_ARG;
    if (!bytes) return LUZRT_OK;
    LUZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

luzrt_status luzrt_download(luzrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return LUZRT_E_INVALID_ARG;
    if (!bytes) return LUZRT_OK;
    if (build_stream_sync(ctx) != LUZRT_OK) return LUZRT_E_DEVICE;   // a TLAS being inspected
    LUZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

// ---------------------------------------------------------------- BLAS
luzrt_status luzrt_blas_build(luzrt_ctx* ctx, const void* vertices, uint32_t stride,
                              uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                              luzrt_blas** out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!out) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_blas_build: null output");
    *out = nullptr;
    if (index_count % 3 != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "index_count %u is not a multiple of 3", index_count);
    if (stride < 12 || stride % 4 != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "vertex stride %u must be >= 12 and a multiple of 4",
                    stride);
    if ((index_count && !indices) || (vertex_count && !vertices))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_blas_build: null buffer");
    // the build replaces an out-of-range index by vertex 0 (and reports it), so
    // triangles need at least one vertex
    if (index_count && vertex_count == 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_blas_build: %u indices but no vertices", index_count);
    if (index_count / 3 >= 0x7FFFFFFFu)
        return fail(ctx, LUZRT_E_LIMIT, "too many triangles");
    LUZ_HIP(ctx, hipSetDevice(ctx->device));
    luzrt_blas* b = new luzrt_blas();
    b->tri_count = index_count / 3;
    b->node_count = b->tri_count > 1 ? b->tri_count - 1 : 1;
    b->vertex_count = vertex_count;
    b->stride = stride;
    const size_t n = b->tri_count ? b->tri_count : 1;
    auto alloc = [&](void** p, size_t bytes) { return hipMalloc(p, bytes); };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = alloc((void**)&b->nodes, (size_t)b->node_count * sizeof(Node));
    if (e == hipSuccess) e = alloc((void**)&b->tris, n * 36);
    if (e == hipSuccess) e = alloc((void**)&b->nrms, n * 36);
    if (e == hipSuccess) e = alloc((void**)&b->prims, n * 4);
    if (e == hipSuccess) e = alloc((void**)&b->slot_idx, n * 12);
    if (e == hipSuccess) e = alloc((void**)&b->leaf_parent, n * 4);
    if (e == hipSuccess) e = alloc((void**)&b->counters, (size_t)b->node_count * 4);
    if (e == hipSuccess) e = alloc((void**)&b->node_end, (size_t)b->node_count * 4);
    if (e == hipSuccess) e = alloc((void**)&b->span, (n + 1) * 4);
    if (e == hipSuccess) e = hipMemset(b->span, 0, 4);   // refit hand-over list starts empty
    if (e != hipSuccess) {
        luzrt_blas_destroy(ctx, b);
        return hip_fail(ctx, e, "hipMalloc(BLAS)");
    }
    // Validate indices on the device-side copy would need a kernel; the
    // caller contract is indices < vertex_count (as for vkCmdBuildAccelerationStructuresKHR).
    stage_begin(ctx, kStageBuild);
    luzrt_status s = blas_build_device(ctx, b, vertices, indices);
    stage_end(ctx, kStageBuild);
    if (s == LUZRT_OK) s = check_build_flag(ctx);
    if (s != LUZRT_OK) {
        luzrt_blas_destroy(ctx, b);
        return s;
    }
    *out = b;
    return LUZRT_OK;
}

luzrt_status luzrt_blas_refit(luzrt_ctx* ctx, luzrt_blas* b, const void* vertices) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b || (b->vertex_count && !vertices))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_blas_refit: null argument");
    // a TLAS build still pending may be reading this BLAS's boxes
    if (ctx->pending_builds > 0 && build_stream_sync(ctx) != LUZRT_OK) return LUZRT_E_DEVICE;
    stage_begin(ctx, kStageBuild);
    luzrt_status s = blas_refit_device(ctx, b, vertices);
    stage_end(ctx, kStageBuild);
    b->wide_valid = false;  // the 8-wide form is rebuilt from the refitted tree on next use
    return s;  // a malformed-tree flag (never set by a refit of a built BLAS) surfaces in luzrt_get_timings
}

luzrt_status luzrt_blas_destroy(luzrt_ctx* ctx, luzrt_blas* b) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b) return LUZRT_OK;
    if (b->refs > 0)
        return fail(ctx, LUZRT_E_IN_USE, "BLAS is referenced by %d TLAS(es)", b->refs);
    hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {b->nodes, b->tris, b->nrms, b->prims, b->slot_idx, b->leaf_parent, b->counters,
                    b->node_end, b->span, b->wide, b->tris8, b->prims8};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete b;
    return LUZRT_OK;
}

luzrt_status luzrt_blas_arrays(const luzrt_blas* b, const void** nodes, uint32_t* node_count,
                               const float** tris, const uint32_t** prims, uint32_t* tri_count) {
    if (!b) return LUZRT_E_INVALID_ARG;
    if (nodes) *nodes = b->nodes;
    if (node_count) *node_count = b->node_count;
    if (tris) *tris = b->tris;
    if (prims) *prims = b->prims;
    if (tri_count) *tri_count = b->tri_count;
    return LUZRT_OK;
}

// ---------------------------------------------------------------- TLAS
luzrt_status luzrt_tlas_create(luzrt_ctx* ctx, uint32_t max_instances, luzrt_tlas** out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!out) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_tlas_create: null output");
    *out = nullptr;
    if (max_instances == 0 || max_instances > LUZRT_MAX_INSTANCES)
        return fail(ctx, LUZRT_E_LIMIT, "max_instances %u outside 1..%d", max_instances,
                    LUZRT_MAX_INSTANCES);
    luzrt_tlas* t = new luzrt_tlas();
    t->capacity = max_instances;
    const size_t n = max_instances;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void**)&t->nodes, n * sizeof(Node));
    if (e == hipSuccess) e = hipMalloc((void**)&t->recs, n * sizeof(InstRec));
    if (e == hipSuccess) e = hipMalloc((void**)&t->aux, n * sizeof(InstAux));
    if (e == hipSuccess) e = hipMalloc((void**)&t->w2o, n * 48);
    if (e == hipSuccess) e = hipMalloc((void**)&t->order, n * 4);
    if (e == hipSuccess) e = hipMalloc(&t->input, n * sizeof(InstIn));
    if (e == hipSuccess) e = hipMalloc(&t->scratch, tlas_scratch_bytes((uint32_t)n) + 256);
    if (e != hipSuccess) {
        luzrt_tlas_destroy(ctx, t);
        return hip_fail(ctx, e, "hipMalloc(TLAS)");
    }
    *out = t;
    return LUZRT_OK;
}

}  // extern "C"

namespace luz {

luzrt_status pack_instances(luzrt_ctx* ctx, luzrt_tlas* t, const luzrt_instance* inst,
                            uint32_t count, std::vector<InstIn>& host) {
    if (!t || (count && !inst)) return fail(ctx, LUZRT_E_INVALID_ARG, "instances: null argument");
    if (count > t->capacity)
        return fail(ctx, LUZRT_E_LIMIT, "%u instances exceed the TLAS capacity %u", count,
                    t->capacity);
    host.clear();
    host.reserve(count);
    std::vector<const luzrt_blas*> refs;
    for (uint32_t i = 0; i < count; ++i) {
        const luzrt_instance& in = inst[i];
        if (!in.blas) return fail(ctx, LUZRT_E_INVALID_ARG, "instance %u has no BLAS", i);
        InstIn r;
        std::memset(&r, 0, sizeof(r));
        std::memcpy(r.o2w, in.object_to_world, 48);
        for (int k = 0; k < 12; ++k)
            if (!std::isfinite(r.o2w[k]))
                return fail(ctx, LUZRT_E_INVALID_ARG, "instance %u transform is not finite", i);
        if (!invert_3x4(r.o2w, r.w2o, r.nmat))
            return fail(ctx, LUZRT_E_INVALID_ARG, "instance %u transform is singular", i);
        r.custom_index = in.custom_index;
        r.mask = in.mask;
        r.input_index = i;
        r.nodes = in.blas->nodes;
        r.tris = in.blas->tris;
        r.nrms = in.blas->nrms;
        r.prims = in.blas->prims;
        if (ctx->precision != LUZRT_PRECISION_PLAIN && in.blas->tri_count > 0) {
            luzrt_blas* b = const_cast<luzrt_blas*>(in.blas);
            if (!b->wide_valid) {
                luzrt_status ws = blas_wide_device(ctx, b);
                if (ws != LUZRT_OK) return ws;
            }
            r.wnodes = b->wide;
            r.wtris = b->tris8;
        }
        refs.push_back(in.blas);
        // empty meshes and mask-0 instances can never be hit by a 0xFF ray mask
        if (in.blas->tri_count == 0 || (in.mask & 0xFFu) == 0) continue;
        host.push_back(r);
    }
    for (const luzrt_blas* b : t->referenced) const_cast<luzrt_blas*>(b)->refs--;
    t->referenced = refs;
    for (const luzrt_blas* b : t->referenced) const_cast<luzrt_blas*>(b)->refs++;
    t->count = (uint32_t)host.size();
    t->node_count = t->count > 1 ? t->count - 1 : 1;
    return LUZRT_OK;
}

}  // namespace luz

extern "C" {

luzrt_status luzrt_tlas_build(luzrt_ctx* ctx, luzrt_tlas* t, const luzrt_instance* inst,
                              uint32_t count) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    std::vector<InstIn> host;
    luzrt_status s = pack_instances(ctx, t, inst, count, host);
    if (s != LUZRT_OK) return s;
    LUZ_HIP(ctx, hipSetDevice(ctx->device));
    t->wide = ctx->precision != LUZRT_PRECISION_PLAIN;   // COMPACT and WIDE trace the 8-wide form
    if (t->wide && !t->recs_w)
        LUZ_HIP(ctx, hipMalloc((void**)&t->recs_w, (size_t)t->capacity * sizeof(InstRec)));
    if (!t->built) LUZ_HIP(ctx, hipEventCreateWithFlags(&t->built, hipEventDisableTiming));
    // On the build stream after the work enqueued so far on the context's
    // stream that reads a TLAS or writes a BLAS (all of it, unless the last
    // such work was a visibility call's traversal: tlas_release); on the
    // context's stream itself while a caller captures it into a graph.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    LUZ_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cs));
    hipStream_t bs = ctx->stream;
    if (cs == hipStreamCaptureStatusNone) {
        bs = ctx->build_stream;
        if (!ctx->release_clean) LUZ_HIP(ctx, hipEventRecord(ctx->build_mark, ctx->stream));
        LUZ_HIP(ctx, hipStreamWaitEvent(bs, ctx->release_clean ? ctx->tlas_release : ctx->build_mark, 0));
    }
    if ((s = stage_upload(ctx, t->input, host.data(), host.size() * sizeof(InstIn), bs)) != LUZRT_OK)
        return s;
    stage_begin(ctx, kStageBuild, bs);
    s = tlas_build_device(ctx, t, (const InstIn*)t->input, (uint32_t)host.size(), bs);
    stage_end(ctx, kStageBuild, bs);
    if (s != LUZRT_OK) return s;
    if (bs != ctx->stream) {
        LUZ_HIP(ctx, hipEventRecord(t->built, bs));
        if (!t->build_pending) ctx->pending_builds++;
        t->build_pending = true;
    }
    return LUZRT_OK;
}

luzrt_status luzrt_tlas_destroy(luzrt_ctx* ctx, luzrt_tlas* t) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!t) return LUZRT_OK;
    hipStreamSynchronize(ctx->stream);
    if (ctx->build_stream) hipStreamSynchronize(ctx->build_stream);
    if (t->built) hipEventDestroy(t->built);
    if (t->build_pending) ctx->pending_builds--;
    for (const luzrt_blas* b : t->referenced) const_cast<luzrt_blas*>(b)->refs--;
    void* ptrs[] = {t->nodes, t->recs, t->recs_w, t->aux, t->w2o, t->order, t->input, t->scratch};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete t;
    return LUZRT_OK;
}

luzrt_status luzrt_tlas_arrays(const luzrt_tlas* t, const void** nodes, uint32_t* node_count,
                               const float** w2o, const uint32_t** order, uint32_t* count) {
    if (!t) return LUZRT_E_INVALID_ARG;
    if (nodes) *nodes = t->nodes;
    if (node_count) *node_count = t->node_count;
    if (w2o) *w2o = t->w2o;
    if (order) *order = t->order;
    if (count) *count = t->count;
    return LUZRT_OK;
}

}  // extern "C"
