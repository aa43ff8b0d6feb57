// shade.hip — the GGX metallic/roughness lighting kernel on gfx950.
//
// Restates light.frag:main (:171-235) with the BRDF of :17-49: G-buffer and
// visibility counts in, RGBA32F out, one thread per pixel reading 36 bytes
// of G-buffer, (lights + 1) count bytes and writing 16 bytes, all coalesced.
// The sRGB decode pow(c/255, 2.2) (light.frag:172) is a 256-entry table.
#include <cmath>

#include <hip/hip_fp16.h>

#include "internal.hpp"
#include "lightpath.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;
constexpr float kPi = kPiF;

__global__ __launch_bounds__(kBlock) void k_shade(const luzrt_gbuffer g, const DevScene* ds,
                                                  const float* lut, const uint8_t* shadow,
                                                  const uint8_t* ao, float4* out) {
    __shared__ float s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.width * g.height) return;
    out[p] = shade_pixel(g, ds, s_lut, CountVisibility{shadow, ao, g.width * g.height}, p);
}

// Light pass without count planes: thread q shades background pixel q and
// active pixel q (active order), the latter's visibility summed from its hit
// bytes (HitVisibility), as k_resolve + k_shade would.
// The read-only inputs are __restrict__: the scene block's lights then come
// through scalar loads (uniform across the wave) instead of vector loads.
__global__ __launch_bounds__(kBlock) void k_shade_hits(const luzrt_gbuffer g, const DevScene* __restrict__ ds,
                                                       const float* __restrict__ lut,
                                                       const uint32_t* __restrict__ act_index,
                                                       const uint32_t* __restrict__ act_list,
                                                       const uint8_t* __restrict__ hits,
                                                       const uint32_t* __restrict__ counters,
                                                       float4* __restrict__ out) {
    __shared__ float s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const uint32_t npix = g.width * g.height;
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= npix) return;
    const uint32_t A = counters[kCntActive];
    if (act_index[q] == kNone) out[q] = shade_pixel(g, ds, s_lut, HitVisibility{hits, ds, 0u, 0u}, q);
    if (q < A) {
        uint32_t first, step;
        ray_row(q, A, ds->r_total, ds->ray_group, first, step);
        const uint32_t p = act_list[q];
        out[p] = shade_pixel(g, ds, s_lut, HitVisibility{hits, ds, first, step}, p);
    }
}

// The same lighting from denoised visibility and AO fractions.
__global__ __launch_bounds__(kBlock) void k_shade_fractions(const luzrt_gbuffer g, const DevScene* ds,
                                                            const float* lut, const float* vis,
                                                            const float* ao, float4* out) {
    __shared__ float s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= g.width * g.height) return;
    out[p] = shade_pixel(g, ds, s_lut, FractionVisibility{vis, ao, g.width * g.height}, p);
}

// Compact form of the lighting kernel: per-light geometry and the specular
// scalar D*G/(4 NdotV NdotL + 1e-4) stay FP32 (the GGX distribution spans
// 1/(pi a^2), beyond FP16 range for smooth surfaces); the Fresnel and diffuse
// colour terms of two lights at a time run in packed FP16 (one half2 lane
// per light), and every light's contribution is accumulated in FP32.
struct LightGeo {
    float spec, NdotL, HV, scale;
};

__device__ inline LightGeo light_geo(const luzrt_light& Lb, V3 P, V3 N, V3 V, float a2, float kk,
                                     float NdotV, float gv, float sh) {
    LightGeo g;
    V3 Lr = vsub(v3(Lb.position[0], Lb.position[1], Lb.position[2]), P);
    V3 L = vnorm(Lr);
    float att = 1.0f;
    V3 ndir = v3(-Lb.direction[0], -Lb.direction[1], -Lb.direction[2]);
    if (Lb.type == LUZRT_LIGHT_DIRECTIONAL) {
        L = vnorm(ndir);
    } else if (Lb.type == LUZRT_LIGHT_SPOT) {
        float dist = vlen(Lr);
        att = 1.0f / (dist * dist);
        float theta = vdot(L, vnorm(ndir));
        float eps = Lb.inner_angle - Lb.outer_angle;
        att = att * clamp01((theta - Lb.outer_angle) / eps);
    } else {
        float dist = vlen(Lr);
        att = 1.0f / (dist * dist);
    }
    V3 H = vnorm(vadd(V, L));
    float NdotH = fmax_(vdot(N, H), 0.0f);
    float den = NdotH * NdotH * (a2 - 1.0f) + 1.0f;
    float D = a2 / (kPi * den * den);
    g.NdotL = fmax_(vdot(N, L), 0.0f);
    float G = (g.NdotL / (g.NdotL * (1.0f - kk) + kk)) * gv;
    g.spec = D * G / (4.0f * NdotV * g.NdotL + 0.0001f);
    g.HV = clamp01(vdot(H, V));
    g.scale = Lb.intensity * att * (1.0f - sh);
    return g;
}

__global__ __launch_bounds__(kBlock) void k_shade_compact(const luzrt_gbuffer g, const DevScene* ds,
                                                          const float* lut, const uint8_t* shadow,
                                                          const uint8_t* ao, float4* out) {
    __shared__ float s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const uint32_t npix = g.width * g.height;
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    const luzrt_scene& S = ds->s;
    const uint32_t px = p % g.width, py = p / g.width;
    V3 N = gb_normal(g, px, py);
    if (vlen(N) == 0.0f) {
        out[p] = make_float4(S.ambient_color[0] * S.ambient_intensity,
                             S.ambient_color[1] * S.ambient_intensity,
                             S.ambient_color[2] * S.ambient_intensity, 1.0f);
        return;
    }
    uint32_t alb = gb_word(g.albedo, g.albedo_pitch, px, py), mat = gb_word(g.material, g.material_pitch, px, py),
             emi = gb_word(g.emission, g.emission_pitch, px, py);
    float albedo[3] = {s_lut[alb & 0xFFu], s_lut[(alb >> 8) & 0xFFu], s_lut[(alb >> 16) & 0xFFu]};
    float rough = (float)(mat & 0xFFu) / 255.0f;
    float metal = (float)((mat >> 8) & 0xFFu) / 255.0f;
    float occl = (float)((mat >> 16) & 0xFFu) / 255.0f;
    V3 P = gb_position(g, px, py);
    V3 V = vnorm(vsub(v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2]), P));
    float a = rough * rough, a2 = a * a;
    float rr = rough + 1.0f, kk = (rr * rr) / 8.0f;
    float NdotV = fmax_(vdot(N, V), 0.0f);
    float gv = NdotV / (NdotV * (1.0f - kk) + kk);
    const __half2 one = __float2half2_rn(1.0f);
    const __half2 hmet = __float2half2_rn(1.0f - metal);
    __half2 F0[3], alb_pi[3];
    for (int c = 0; c < 3; ++c) {
        F0[c] = __float2half2_rn(0.04f * (1.0f - metal) + albedo[c] * metal);
        alb_pi[c] = __float2half2_rn(albedo[c] / kPi);
    }
    float Lo[3] = {0.0f, 0.0f, 0.0f};
    const int nl = S.num_lights;
    for (int l = 0; l < nl; l += 2) {
        const bool pair = l + 1 < nl;
        int n0 = S.lights[l].num_shadow_samples;
        float sh0 = !shadow_from_samples(S, l) ? 1.0f
                    : (n0 > 0 ? (float)shadow[(size_t)l * npix + p] / (float)n0 : 0.0f);
        LightGeo g0 = light_geo(S.lights[l], P, N, V, a2, kk, NdotV, gv, sh0);
        LightGeo g1 = g0;
        if (pair) {
            int n1 = S.lights[l + 1].num_shadow_samples;
            float sh1 = !shadow_from_samples(S, l + 1) ? 1.0f
                        : (n1 > 0 ? (float)shadow[(size_t)(l + 1) * npix + p] / (float)n1 : 0.0f);
            g1 = light_geo(S.lights[l + 1], P, N, V, a2, kk, NdotV, gv, sh1);
        }
        // packed FP16, lane x = light l, lane y = light l+1
        __half2 om = __hsub2(one, __floats2half2_rn(g0.HV, g1.HV));
        __half2 om2 = __hmul2(om, om);
        __half2 om5 = __hmul2(__hmul2(om2, om2), om);
        for (int c = 0; c < 3; ++c) {
            __half2 F = __hadd2(F0[c], __hmul2(__hsub2(one, F0[c]), om5));
            __half2 kd = __hmul2(__hmul2(__hsub2(one, F), hmet), alb_pi[c]);
            float2 f = __half22float2(F), d = __half22float2(kd);
            Lo[c] += (d.x + g0.spec * f.x) * (S.lights[l].color[c] * g0.scale) * g0.NdotL;
            if (pair) Lo[c] += (d.y + g1.spec * f.y) * (S.lights[l + 1].color[c] * g1.scale) * g1.NdotL;
        }
    }
    int na = S.ao_num_samples;
    float aof = na > 0 ? (float)ao[p] / (float)na : 1.0f;
    float o[3];
    for (int k = 0; k < 3; ++k) {
        float amb = S.ambient_color[k] * S.ambient_intensity * albedo[k] * occl * aof;
        float e = (float)((emi >> (8 * k)) & 0xFFu) / 255.0f;
        o[k] = amb + Lo[k] + e;
    }
    out[p] = make_float4(o[0], o[1], o[2], 1.0f);
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

luzrt_status ensure_lut(luzrt_ctx* ctx) {
    if (ctx->lut.ptr) return LUZRT_OK;
    luzrt_status s = ensure(ctx, ctx->lut, 256 * sizeof(float));
    if (s != LUZRT_OK) return s;
    float h[256];
    for (int i = 0; i < 256; ++i) h[i] = std::pow((float)i / 255.0f, 2.2f);
    LUZ_HIP(ctx, hipMemcpyAsync(ctx->lut.ptr, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

// g: a checked descriptor (check_gbuffer).
luzrt_status shade_device(luzrt_ctx* ctx, const luzrt_gbuffer* g, const uint8_t* shadow,
                          const uint8_t* ao, float* out) {
    luzrt_status s = ensure_lut(ctx);
    if (s != LUZRT_OK) return s;
    const size_t npix = (size_t)g->width * g->height;
    stage_begin(ctx, kStageShade);
    if (ctx->precision == LUZRT_PRECISION_COMPACT)
        k_shade_compact<<<cdiv(npix, kBlock), kBlock, 0, ctx->stream>>>(
            *g, (const DevScene*)ctx->scene.ptr, (const float*)ctx->lut.ptr, shadow, ao, (float4*)out);
    else
        k_shade<<<cdiv(npix, kBlock), kBlock, 0, ctx->stream>>>(
            *g, (const DevScene*)ctx->scene.ptr, (const float*)ctx->lut.ptr, shadow, ao, (float4*)out);
    stage_end(ctx, kStageShade);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

}  // namespace luz

using namespace luz;

extern "C" luzrt_status luzrt_shade(luzrt_ctx* ctx, const luzrt_gbuffer* g, const luzrt_scene* scene,
                                    const uint8_t* shadow, const uint8_t* ao, float* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!g || !scene || !ao || !out || (scene->num_lights > 0 && !shadow))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shade: null argument");
    if (((uintptr_t)out & 15u) != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shade: output must be 16-byte aligned");
    DevScene ds;
    luzrt_status s = upload_scene(ctx, scene, &ds);
    if (s != LUZRT_OK) return s;
    luzrt_gbuffer gl;
    if ((s = check_gbuffer(ctx, g, &gl)) != LUZRT_OK) return s;
    return shade_device(ctx, &gl, shadow, ao, out);
}

extern "C" luzrt_status luzrt_shade_fractions(luzrt_ctx* ctx, const luzrt_gbuffer* g,
                                              const luzrt_scene* scene, const float* visibility,
                                              const float* ao, float* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!g || !scene || !ao || !out || (scene->num_lights > 0 && !visibility))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shade_fractions: null argument");
    if (((uintptr_t)out & 15u) != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_shade_fractions: output must be 16-byte aligned");
    DevScene ds;
    luzrt_status s = upload_scene(ctx, scene, &ds);
    if (s != LUZRT_OK) return s;
    luzrt_gbuffer gl;
    if ((s = check_gbuffer(ctx, g, &gl)) != LUZRT_OK) return s;
    if ((s = ensure_lut(ctx)) != LUZRT_OK) return s;
    const uint32_t npix = g->width * g->height;
    stage_begin(ctx, kStageShade);
    k_shade_fractions<<<(npix + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(
        gl, (const DevScene*)ctx->scene.ptr, (const float*)ctx->lut.ptr, visibility, ao, (float4*)out);
    stage_end(ctx, kStageShade);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

extern "C" luzrt_status luzrt_light_pass(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                         const luzrt_gbuffer* g, const luzrt_scene* scene,
                                         float* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !g || !scene || !out)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_light_pass: null argument");
    if (((uintptr_t)out & 15u) != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_light_pass: output must be 16-byte aligned");
    luzrt_gbuffer gl;
    luzrt_status s = check_gbuffer(ctx, g, &gl);
    if (s != LUZRT_OK) return s;
    const size_t npix = (size_t)g->width * g->height;
    if (ctx->precision != LUZRT_PRECISION_COMPACT) {
        // visibility without count planes, then lighting straight from the hit bytes
        if ((s = visibility_core(ctx, tlas, g, scene, nullptr, nullptr, false)) != LUZRT_OK) return s;
        if ((s = ensure_lut(ctx)) != LUZRT_OK) return s;
        stage_begin(ctx, kStageShade);
        k_shade_hits<<<cdiv(npix, kBlock), kBlock, 0, ctx->stream>>>(
            gl, (const DevScene*)ctx->scene.ptr, (const float*)ctx->lut.ptr,
            (const uint32_t*)ctx->act_index.ptr, (const uint32_t*)ctx->active.ptr,
            (const uint8_t*)ctx->hits.ptr, (const uint32_t*)ctx->counters.ptr, (float4*)out);
        stage_end(ctx, kStageShade);
        LUZ_HIP(ctx, hipGetLastError());
        return LUZRT_OK;
    }
    int nl = scene->num_lights < 0 ? 0 : scene->num_lights;
    if ((s = ensure(ctx, ctx->vis_shadow, npix * (size_t)(nl > 0 ? nl : 1))) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->vis_ao, npix)) != LUZRT_OK) return s;
    s = luzrt_trace_visibility(ctx, tlas, g, scene, (uint8_t*)ctx->vis_shadow.ptr,
                               (uint8_t*)ctx->vis_ao.ptr);
    if (s != LUZRT_OK) return s;
    return shade_device(ctx, &gl, (const uint8_t*)ctx->vis_shadow.ptr, (const uint8_t*)ctx->vis_ao.ptr,
                        out);
}
