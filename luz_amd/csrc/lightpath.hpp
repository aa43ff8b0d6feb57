// lightpath.hpp — per-pixel stages of the lighting path, shared by the
// single-scene kernels (trace.hip, shade.hip) and the batched multi-scene
// kernels (batch.hip), so both evaluate identical arithmetic.
//
//   tile_pixel / pixel_active / block_prefix  compaction of non-background pixels
//   shadow_setup / gen_ray                    light.frag:86-146, 111-135, 229-231
//   resolve_pixel                             occluded / unoccluded sample counts
//   shade_pixel                               light.frag:17-49, 171-235 (FP32, hardware rcp/rsq)
//   primary_pixel / primary_texel             G-buffer texels by closest-hit primary rays
#pragma once
#include "internal.hpp"
#include "traverse.hpp"

namespace luz {

constexpr float kPiF = 3.14159265359f;  // LuzCommon.h:10

__device__ inline float clamp01(float a) { return a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a); }

// ------------------------------------------------------------ G-buffer texels
// A luzrt_gbuffer reaches the device with its layout filled in (the host's
// check_gbuffer replaces zero strides and pitches by the packed defaults), so
// texel (x, y) of a plane is at base + y * pitch + x * stride.
__device__ inline const float* gb_f3(const float* base, uint32_t pitch, uint32_t stride, uint32_t x,
                                     uint32_t y) {
    return reinterpret_cast<const float*>(reinterpret_cast<const uint8_t*>(base) + (size_t)y * pitch +
                                          (size_t)x * stride);
}

__device__ inline uint32_t gb_word(const uint32_t* base, uint32_t pitch, uint32_t x, uint32_t y) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(base) + (size_t)y * pitch +
                                              (size_t)x * 4u);
}

__device__ inline V3 gb_normal(const luzrt_gbuffer& g, uint32_t x, uint32_t y) {
    const float* n = gb_f3(g.normal, g.normal_pitch, g.normal_stride, x, y);
    return v3(n[0], n[1], n[2]);
}

// out[r] = sum over c of m[c][r] v[c] (column-major), summed in column order.
__device__ inline void mat4_mul(const float* m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r) {
        float a = m[r] * v[0];
        a = a + m[4 + r] * v[1];
        a = a + m[8 + r] * v[2];
        a = a + m[12 + r] * v[3];
        out[r] = a;
    }
}

// DepthToWorld (utils.glsl:1-7) at the centre of pixel (x, y): the clip-space
// point (uv * 2 - 1, depth, 1) through the inverse projection, divided by its
// w (correctly rounded), then through the inverse view.
__device__ inline V3 depth_to_world(const luzrt_gbuffer& g, uint32_t x, uint32_t y) {
    const float d = gb_f3(g.depth, g.depth_pitch, 4u, x, y)[0];
    const float u = __fdiv_rn((float)x + 0.5f, (float)g.width);
    const float v = __fdiv_rn((float)y + 0.5f, (float)g.height);
    const float clip[4] = {u * 2.0f - 1.0f, v * 2.0f - 1.0f, d, 1.0f};
    float vs[4], ws[4];
    mat4_mul(g.inverse_proj, clip, vs);
    const float w = vs[3];
    for (int k = 0; k < 4; ++k) vs[k] = __fdiv_rn(vs[k], w);
    mat4_mul(g.inverse_view, vs, ws);
    return v3(ws[0], ws[1], ws[2]);
}

__device__ inline V3 gb_position(const luzrt_gbuffer& g, uint32_t x, uint32_t y) {
    if (!g.position) return depth_to_world(g, x, y);
    const float* q = gb_f3(g.position, g.position_pitch, g.position_stride, x, y);
    return v3(q[0], q[1], q[2]);
}

// Pixel g of the 64-pixel-tile order with tiles tw wide and 64/tw tall.
__device__ inline void tile_pixel(uint32_t g, uint32_t W, uint32_t tw, uint32_t& x, uint32_t& y) {
    uint32_t tile = g >> 6, t = g & 63u;
    uint32_t th = 64u / tw;
    uint32_t tiles_x = (W + tw - 1) / tw;
    x = (tile % tiles_x) * tw + (t % tw);
    y = (tile / tiles_x) * th + t / tw;
}

// Tile-order pixel g of a W x H G-buffer: its index p and whether it is
// inside the image and not background (non-zero normal, light.frag:178).
__device__ inline bool pixel_active(const float* normal, uint32_t g, uint32_t W, uint32_t H,
                                    uint32_t tw, uint32_t ntiles_px, uint32_t& p, bool& inside) {
    uint32_t x = 0, y = 0;
    p = 0;
    inside = false;
    if (g >= ntiles_px) return false;
    tile_pixel(g, W, tw, x, y);
    if (x >= W || y >= H) return false;
    inside = true;
    p = y * W + x;
    const float* n = normal + (size_t)p * 3;
    return vlen(v3(n[0], n[1], n[2])) != 0.0f;
}

// As above, through the layout of a G-buffer descriptor.
__device__ inline bool pixel_active(const luzrt_gbuffer& gb, uint32_t g, uint32_t tw, uint32_t ntiles_px,
                                    uint32_t& p, bool& inside) {
    uint32_t x = 0, y = 0;
    p = 0;
    inside = false;
    if (g >= ntiles_px) return false;
    tile_pixel(g, gb.width, tw, x, y);
    if (x >= gb.width || y >= gb.height) return false;
    inside = true;
    p = y * gb.width + x;
    return vlen(gb_normal(gb, x, y)) != 0.0f;
}

// Exclusive rank of this thread's flag within its 256-thread block; the
// block total in block_total.  All threads of the block must call it.
__device__ inline uint32_t block_prefix(bool active, uint32_t* wave_tot, uint32_t& block_total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(active);
    uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
        uint32_t c = wave_tot[w];
        if (w < wave) off += c;
        tot += c;
    }
    block_total = tot;
    return off + (uint32_t)__popcll(m & lt);
}

struct ShadowSetup { V3 Lvec, T, B; float tmax; };

// light.frag:137-146 (origin, ray vector) and :90-97 (tangent frame, tMax).
__device__ inline ShadowSetup shadow_setup(const luzrt_light& L, V3 P) {
    ShadowSetup r;
    V3 lpos = v3(L.position[0], L.position[1], L.position[2]);
    V3 Lraw = vsub(lpos, P);
    float dist = vlen(Lraw);
    if (L.type == LUZRT_LIGHT_DIRECTIONAL) {
        V3 dir = v3(L.direction[0], L.direction[1], L.direction[2]);
        V3 Ld = vnorm(v3(-dir.x, -dir.y, -dir.z));
        r.Lvec = vmul(vmul(dir, vdot(dir, Ld)), dist);
    } else {
        r.Lvec = vmul(vnorm(Lraw), dist);
    }
    V3 c = vcross(r.Lvec, v3(0.0f, 1.0f, 0.0f));
    if (vlen(c) == 0.0f) c = vcross(r.Lvec, v3(1.0f, 0.0f, 0.0f));
    r.T = vnorm(c);
    r.B = vnorm(vcross(r.T, r.Lvec));
    r.tmax = vlen(r.Lvec);
    return r;
}

// A G-buffer texel as the ray generators see it: pixel coordinates, position,
// normal and distance to the camera.
struct PixelFrame {
    uint32_t x, y;
    V3 P, N;
    float lc;
};

__device__ inline PixelFrame pixel_frame(const float* position, const float* normal, uint32_t W,
                                         const luzrt_scene& S, uint32_t p) {
    PixelFrame f;
    f.x = p % W;
    f.y = p / W;
    f.P = v3(position[(size_t)p * 3], position[(size_t)p * 3 + 1], position[(size_t)p * 3 + 2]);
    f.N = v3(normal[(size_t)p * 3], normal[(size_t)p * 3 + 1], normal[(size_t)p * 3 + 2]);
    f.lc = vlen(vsub(f.P, v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2])));
    return f;
}

// The frame of pixel p through the layout of a G-buffer descriptor.
__device__ inline PixelFrame pixel_frame(const luzrt_gbuffer& g, const luzrt_scene& S, uint32_t p) {
    PixelFrame f;
    f.x = p % g.width;
    f.y = p / g.width;
    f.P = gb_position(g, f.x, f.y);
    f.N = gb_normal(g, f.x, f.y);
    f.lc = vlen(vsub(f.P, v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2])));
    return f;
}

// The frame of an active pixel as the compaction stores it for the traversal
// (two 16-byte rows: position and camera distance; normal and x | y << 16),
// so that each ray's generation reads one record instead of the active list,
// then the position and normal planes, and does not recompute the distance.
__device__ inline void pixel_record(const PixelFrame& f, float4& r0, float4& r1) {
    r0 = make_float4(f.P.x, f.P.y, f.P.z, f.lc);
    r1 = make_float4(f.N.x, f.N.y, f.N.z, __uint_as_float(f.x | (f.y << 16)));
}

__device__ inline PixelFrame pixel_from_record(float4 r0, float4 r1) {
    PixelFrame f;
    const uint32_t xy = __float_as_uint(r1.w);
    f.x = xy & 0xFFFFu;
    f.y = xy >> 16;
    f.P = v3(r0.x, r0.y, r0.z);
    f.N = v3(r1.x, r1.y, r1.z);
    f.lc = r0.w;
    return f;
}

// Shadow-ray origin: P lifted along N by max(0.01 x camera distance, 0.05)
// (light.frag:137-146).
__device__ inline V3 shadow_origin(const PixelFrame& f) {
    const float sb = fmax_(f.lc * 0.01f, 0.05f);
    return vadd(f.P, vmul(f.N, sb));
}

// Direction of shadow sample s of light l: the light vector offset by a
// disk sample of the light's radius in its tangent frame (light.frag:111-135).
__device__ inline V3 shadow_dir(const luzrt_light& L, const ShadowSetup& ss, const PixelFrame& f,
                                uint32_t frame, int l, int s) {
    float u, v;
    sample2(f.x, f.y, frame, (uint32_t)(l * 256 + s), u, v);
    const float pr = L.radius * sqrtf(u);
    float sn, cs;
    sincos_turns(v, sn, cs);
    const float dx = pr * cs, dy = pr * sn;
    const V3 d = v3(ss.Lvec.x + dx * ss.T.x + dy * ss.B.x,
                    ss.Lvec.y + dx * ss.T.y + dy * ss.B.y,
                    ss.Lvec.z + dx * ss.T.z + dy * ss.B.z);
    return vnorm(d);
}

// AO-ray origin: P lifted along N by 0.01 x camera distance (light.frag:229-231).
__device__ inline V3 ao_origin(const PixelFrame& f) {
    const float ab = f.lc * 0.01f;
    return vadd(f.P, vmul(f.N, ab));
}

// Cosine-weighted hemisphere sample s around N (light.frag:86-97, 118-126).
__device__ inline V3 ao_dir(const PixelFrame& f, uint32_t frame, int s) {
    const V3 N = f.N;
    const V3 T = fabsf(N.z) > 0.5f ? v3(0.0f, -N.z, N.y) : v3(-N.y, N.x, 0.0f);
    const V3 B = vcross(N, T);
    float u, v;
    sample2(f.x, f.y, frame, (uint32_t)(16384 + s), u, v);
    const float r = sqrtf(u);
    float sn, cs;
    sincos_turns(v * kHemiTurns, sn, cs);
    const float hx = r * cs, hy = r * sn;
    const float om = 1.0f - u;
    const float hz = sqrtf(om > 0.0f ? om : 0.0f);
    return v3(T.x * hx + B.x * hy + N.x * hz,
              T.y * hx + B.y * hy + N.y * hz,
              T.z * hx + B.z * hy + N.z * hz);
}

// The light that owns shadow slot j < r_shadow: the last light l with
// slot_base[l] <= j (lights without samples own no slots).
__device__ inline int slot_light(const DevScene* ds, uint32_t j) {
    int lo = 0, hi = ds->s.num_lights;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ds->slot_base[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

// The ray of slot j from the frame f of an active pixel, slot j < r_shadow
// belonging to light l = slot_light(j) (l is ignored for AO slots).  Slot
// j < r_shadow is shadow sample j - slot_base[l] of light l; slot j >= r_shadow
// is AO sample j - r_shadow.  Shadow rays start at shadow_origin with tmin
// 0.001 and end at the light (tmax = |light vector|); AO rays span
// [ao_min, ao_max] from ao_origin.
__device__ inline void gen_ray_of(const DevScene* ds, const PixelFrame& f, uint32_t j, int l, V3& o,
                                  V3& d, float& tmin, float& tmax) {
    const luzrt_scene& S = ds->s;
    const uint32_t frame = (uint32_t)S.frame;
    if (j < ds->r_shadow) {
        const luzrt_light& L = S.lights[l];
        const ShadowSetup ss = shadow_setup(L, f.P);
        o = shadow_origin(f);
        d = shadow_dir(L, ss, f, frame, l, (int)(j - ds->slot_base[l]));
        tmin = 0.001f;
        tmax = ss.tmax;
    } else {
        o = ao_origin(f);
        d = ao_dir(f, frame, (int)(j - ds->r_shadow));
        tmin = S.ao_min;
        tmax = S.ao_max;
    }
}

__device__ inline void gen_ray(const DevScene* ds, const PixelFrame& f, uint32_t j, V3& o, V3& d,
                               float& tmin, float& tmax) {
    gen_ray_of(ds, f, j, j < ds->r_shadow ? slot_light(ds, j) : 0, o, d, tmin, tmax);
}

// Ray order of a visibility launch over A active pixels with R slots each:
// groups of G consecutive active pixels (G = DevScene::ray_group, or one group
// of all A pixels when that is 0), slot-major inside a group, so ray
// (g * R + j) * G + i is slot j of active pixel g * G + i (the last, partial
// group likewise with its own width).  A wave's 64 lanes hold one slot (one
// light sample or AO sample) of 64 neighbouring pixels either way; with small
// groups the R rays of a pixel are also claimed close together, so its
// G-buffer texel comes from cache after the first of them.
//
// ray_row: the ray of slot 0 of active pixel a and the step between its slots.
__device__ inline void ray_row(uint32_t a, uint32_t A, uint32_t R, uint32_t G, uint32_t& first,
                               uint32_t& step) {
    if (G == 0) {
        first = a;
        step = A;
        return;
    }
    const uint32_t g = a / G, full = A / G;
    step = g < full ? G : A - full * G;
    first = g * G * R + (a - g * G);
}

// Inverse: slot j and active pixel a of ray rid < A * R.
__device__ inline void ray_slot(uint32_t rid, uint32_t A, uint32_t R, uint32_t G, uint32_t& j,
                                uint32_t& a) {
    if (G == 0) {
        j = rid / A;
        a = rid - j * A;
        return;
    }
    const uint32_t g = rid / (G * R), full = A / G;
    const uint32_t w = g < full ? G : A - full * G;
    const uint32_t within = rid - g * G * R;
    j = within / w;
    a = g * G + (within - j * w);
}

// Samples n of a pixel's hit mask from slot j0 (0 < n, j0 + n <= 64).
__device__ inline uint64_t slot_bits(uint64_t m, uint32_t j0, uint32_t n) {
    return (m >> j0) & (n >= 64 ? ~0ull : (1ull << n) - 1ull);
}

// Per-pixel counts from hit bytes: occluded shadow samples per light
// (plane l of `shadow`), unoccluded AO samples; a = kNone for background.
// Up to 64 rays per pixel the hit bytes (0 or 1) are gathered into a bit mask
// first, the loads of 32 slots in flight together, and each count is a
// population count of its slots.
__device__ inline void resolve_pixel(const DevScene* ds, uint32_t a, uint32_t A, const uint8_t* hits,
                                     uint32_t npix, uint32_t p, uint8_t* shadow, uint8_t* ao) {
    const luzrt_scene& S = ds->s;
    const uint32_t R = ds->r_total;
    if (a != kNone && R > 0 && R <= 64) {
        uint32_t first, step;
        ray_row(a, A, R, ds->ray_group, first, step);
        uint64_t m = 0;
        for (uint32_t c = 0; c < R; c += 32) {
            uint32_t w = 0;
#pragma unroll
            for (uint32_t u = 0; u < 32; ++u) {
                const uint32_t j = c + u < R ? c + u : R - 1;   // slots past R re-read the last
                w |= (uint32_t)(hits[first + j * step] != 0) << u;
            }
            if (R - c < 32) w &= (1u << (R - c)) - 1u;
            m |= (uint64_t)w << c;
        }
        for (int l = 0; l < S.num_lights; ++l) {
            const int n = S.lights[l].num_shadow_samples;
            const uint32_t cnt = n > 0 ? (uint32_t)__popcll(slot_bits(m, ds->slot_base[l], (uint32_t)n)) : 0u;
            shadow[(size_t)l * npix + p] = (uint8_t)cnt;
        }
        const int na = S.ao_num_samples;
        ao[p] = (uint8_t)(na > 0 ? (uint32_t)na - (uint32_t)__popcll(slot_bits(m, ds->r_shadow, (uint32_t)na)) : 0u);
        return;
    }
    uint32_t first = 0, step = 0;
    if (a != kNone) ray_row(a, A, ds->r_total, ds->ray_group, first, step);
    for (int l = 0; l < S.num_lights; ++l) {
        uint32_t cnt = 0;
        int n = S.lights[l].num_shadow_samples;
        if (a != kNone && n > 0) {
            uint32_t j0 = ds->slot_base[l];
            for (int s = 0; s < n; ++s) cnt += hits[first + (j0 + s) * step];
        }
        shadow[(size_t)l * npix + p] = (uint8_t)cnt;
    }
    uint32_t miss = 0;
    if (a != kNone) {
        uint32_t j0 = ds->r_shadow;
        for (int s = 0; s < S.ao_num_samples; ++s) miss += 1u - hits[first + (j0 + s) * step];
    }
    ao[p] = (uint8_t)miss;
}

// Visibility from per-pixel sample counts: occluded shadow samples of light
// l (plane l) out of its sample count, unoccluded AO samples out of n_AO
// (light.frag:104-108, 130-134).
struct CountVisibility {
    const uint8_t* shadow;
    const uint8_t* ao;
    uint32_t npix;
    __device__ float shadow_frac(const luzrt_scene& S, uint32_t l, uint32_t p) const {
        int ns = S.lights[l].num_shadow_samples;
        return ns > 0 ? (float)shadow[(size_t)l * npix + p] / (float)ns : 0.0f;
    }
    __device__ float ao_frac(const luzrt_scene& S, uint32_t p) const {
        int na = S.ao_num_samples;
        return na > 0 ? (float)ao[p] / (float)na : 1.0f;
    }
};

// Visibility straight from the hit bytes of one active pixel (its slot-0 ray
// `first` and the step between its slots, ray_row): the same fractions as
// CountVisibility over the counts k_resolve would have written.
// cnt / n for 0 <= cnt <= n, n > 0: exactly 0 and 1 at the ends, else the
// product with the hardware reciprocal (within 2 ulps of the quotient; the
// radiance it scales is compared to 1e-3).
__device__ inline float count_fraction(uint32_t cnt, uint32_t n) {
    if (cnt == 0u) return 0.0f;
    if (cnt == n) return 1.0f;
    return (float)cnt * __builtin_amdgcn_rcpf((float)n);
}

struct HitVisibility {
    const uint8_t* hits;
    const DevScene* ds;
    uint32_t first, step;
    __device__ float shadow_frac(const luzrt_scene& S, uint32_t l, uint32_t) const {
        const int ns = S.lights[l].num_shadow_samples;
        if (ns <= 0) return 0.0f;
        const uint32_t j0 = ds->slot_base[l];
        uint32_t cnt = 0;
        for (int s = 0; s < ns; ++s) cnt += hits[first + (j0 + s) * step];
        return count_fraction(cnt, (uint32_t)ns);
    }
    __device__ float ao_frac(const luzrt_scene& S, uint32_t) const {
        const int na = S.ao_num_samples;
        if (na <= 0) return 1.0f;
        const uint32_t j0 = ds->r_shadow;
        uint32_t miss = 0;
        for (int s = 0; s < na; ++s) miss += 1u - hits[first + (j0 + s) * step];
        return count_fraction(miss, (uint32_t)na);
    }
};

// Visibility from denoised fractions: visibility of light l (plane l) and
// the AO fraction; lights without shadow samples are unshadowed, and AO is
// 1 when the scene casts no AO rays, as with counts.
struct FractionVisibility {
    const float* vis;
    const float* ao;
    uint32_t npix;
    __device__ float shadow_frac(const luzrt_scene& S, uint32_t l, uint32_t p) const {
        return S.lights[l].num_shadow_samples > 0 ? 1.0f - vis[(size_t)l * npix + p] : 0.0f;
    }
    __device__ float ao_frac(const luzrt_scene& S, uint32_t p) const {
        return S.ao_num_samples > 0 ? ao[p] : 1.0f;
    }
};

// EvaluateShadow's result by shadow type (light.frag:141-168): ray-traced
// scenes take the traced fraction (0 without samples, TraceShadowRay
// :87-89); shadow-map scenes take the map decision supplied as one sample
// per light and are fully shadowed for a light without one (shadowMap ==
// -1); every other type (ShadowDisabled) is fully shadowed.  Visibility
// calls zero the sample counts of non-ray-traced scenes, so their light pass
// is fully shadowed as the reference's is.
__device__ inline bool shadow_from_samples(const luzrt_scene& S, uint32_t l) {
    return S.shadow_type == LUZRT_SHADOW_RAYTRACED ||
           (S.shadow_type == LUZRT_SHADOW_MAP && S.lights[l].num_shadow_samples > 0);
}

template <typename VIS>
__device__ inline float shadow_factor(const luzrt_scene& S, const VIS& vis, uint32_t l, uint32_t p) {
    return shadow_from_samples(S, l) ? vis.shadow_frac(S, l, p) : 1.0f;
}

// The lighting arithmetic is compared with the oracle to 1e-3 relative, not
// bit for bit (only visibility decisions are), so its well-conditioned terms
// use the hardware reciprocal, reciprocal square root and square root (about
// 1 ulp each).  The GGX distribution is not well conditioned: near a
// highlight of a smooth surface den = NdotH^2 (a^2 - 1) + 1 cancels to ~a^2
// (5e-5 at roughness 0.09), so a 1-ulp change of |H| moves D by ~0.5%.  The
// chain V, L, H, NdotH, den, D is therefore evaluated with correctly rounded
// square roots and divisions in the oracle's order, and matches it bit for bit.
__device__ inline float rcp_hw(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ inline float sqrt_hw(float a) { return __builtin_amdgcn_sqrtf(a); }
constexpr float kInvPiF = 0.318309886183791f;

// light.frag:main in FP32; lut = pow(i/255, 2.2).
template <typename VIS>
__device__ inline float4 shade_pixel(const luzrt_gbuffer& g, const DevScene* ds, const float* lut,
                                     const VIS& vis, uint32_t p) {
    const luzrt_scene& S = ds->s;
    const uint32_t px = p % g.width, py = p / g.width;
    V3 N = gb_normal(g, px, py);
    if (vlen(N) == 0.0f)
        return make_float4(S.ambient_color[0] * S.ambient_intensity,
                           S.ambient_color[1] * S.ambient_intensity,
                           S.ambient_color[2] * S.ambient_intensity, 1.0f);
    uint32_t alb = gb_word(g.albedo, g.albedo_pitch, px, py), mat = gb_word(g.material, g.material_pitch, px, py),
             emi = gb_word(g.emission, g.emission_pitch, px, py);
    float albedo[3] = {lut[alb & 0xFFu], lut[(alb >> 8) & 0xFFu], lut[(alb >> 16) & 0xFFu]};
    const float inv255 = rcp_hw(255.0f);
    float rough = __fdiv_rn((float)(mat & 0xFFu), 255.0f);   // a^2 enters D: exact, as the oracle
    float metal = (float)((mat >> 8) & 0xFFu) * inv255;
    float occl = (float)((mat >> 16) & 0xFFu) * inv255;
    V3 P = gb_position(g, px, py);
    V3 cam = v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2]);
    V3 V = vnorm(vsub(cam, P));
    float F0[3];
    for (int k = 0; k < 3; ++k) F0[k] = 0.04f * (1.0f - metal) + albedo[k] * metal;
    float Lo[3] = {0.0f, 0.0f, 0.0f};
    float a = rough * rough, a2 = a * a;
    float rr = rough + 1.0f, kk = (rr * rr) / 8.0f;
    float NdotV = fmax_(vdot(N, V), 0.0f);
    float gv = NdotV * rcp_hw(NdotV * (1.0f - kk) + kk);
    const uint32_t nl = (uint32_t)S.num_lights;
    for (uint32_t l = 0; l < nl; ++l) {
        const luzrt_light& Lb = S.lights[l];
        V3 Lr = vsub(v3(Lb.position[0], Lb.position[1], Lb.position[2]), P);
        V3 ndir = v3(-Lb.direction[0], -Lb.direction[1], -Lb.direction[2]);
        V3 L = vnorm(Lb.type == LUZRT_LIGHT_DIRECTIONAL ? ndir : Lr);
        float att = 1.0f;
        if (Lb.type == LUZRT_LIGHT_DIRECTIONAL) {
        } else if (Lb.type == LUZRT_LIGHT_SPOT) {
            att = rcp_hw(vdot(Lr, Lr));
            float theta = vdot(L, vnorm(ndir));
            float eps = Lb.inner_angle - Lb.outer_angle;
            att = att * clamp01(__fdiv_rn(theta - Lb.outer_angle, eps));
        } else {
            att = rcp_hw(vdot(Lr, Lr));
        }
        float sh = shadow_factor(S, vis, l, p);
        float scale = Lb.intensity * att * (1.0f - sh);
        V3 H = vnorm(vadd(V, L));
        float NdotH = fmax_(vdot(N, H), 0.0f);
        float den = NdotH * NdotH * (a2 - 1.0f) + 1.0f;
        float D = __fdiv_rn(a2, kPiF * den * den);
        float NdotL = fmax_(vdot(N, L), 0.0f);
        float G = (NdotL * rcp_hw(NdotL * (1.0f - kk) + kk)) * gv;
        float HV = clamp01(vdot(H, V));
        float om = clamp01(1.0f - HV);
        float om5 = om * om * om * om * om;
        float inv_denom = rcp_hw(4.0f * NdotV * NdotL + 0.0001f);
        float DG = D * G;
        for (int k = 0; k < 3; ++k) {
            float F = F0[k] + (1.0f - F0[k]) * om5;
            float spec = DG * F * inv_denom;
            float kD = (1.0f - F) * (1.0f - metal);
            float radiance = Lb.color[k] * scale;
            Lo[k] = Lo[k] + (kD * albedo[k] * kInvPiF + spec) * radiance * NdotL;
        }
    }
    float aof = vis.ao_frac(S, p);
    float o[3];
    for (int k = 0; k < 3; ++k) {
        float amb = S.ambient_color[k] * S.ambient_intensity * albedo[k] * occl * aof;
        float e = (float)((emi >> (8 * k)) & 0xFFu) * inv255;
        o[k] = amb + Lo[k] + e;
    }
    return make_float4(o[0], o[1], o[2], 1.0f);
}

// G-buffer texel (x, y) from the closest hit of its primary ray: world
// position, normalised interpolated vertex normal through the instance's
// inverse-transpose, and the material of the hit instance (background: all
// zero).  The stack must be empty; it is empty again on return.
// Primary ray of pixel (x, y) through its centre (opaque.vert's camera,
// y flipped as the Vulkan projection does).
__device__ inline V3 primary_dir(const luzrt_camera& cam, uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
    float nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
    float ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
    return v3(cam.forward[0] + nx * cam.right[0] + ny * cam.up[0],
              cam.forward[1] + nx * cam.right[1] + ny * cam.up[1],
              cam.forward[2] + nx * cam.right[2] + ny * cam.up[2]);
}

// G-buffer texel (x, y) from the closest hit `best` of the primary ray `d`
// (opaque.frag:21-58): world position, normalised interpolated vertex normal
// through the instance's inverse-transpose, the instance's material; zeros on
// a miss (background).
__device__ inline void primary_texel(const luzrt_camera& cam, V3 d, const Best& best, const InstAux* aux,
                                     const luzrt_material* mats, uint32_t nmats,
                                     const luzrt_gbuffer_out& out, uint32_t x, uint32_t y) {
    const size_t p = (size_t)y * out.width + x;
    if (!best.found) {
        for (int k = 0; k < 3; ++k) { out.position[p * 3 + k] = 0.0f; out.normal[p * 3 + k] = 0.0f; }
        out.albedo[p] = 0; out.material[p] = 0; out.emission[p] = 0;
        return;
    }
    const V3 eye = v3(cam.eye[0], cam.eye[1], cam.eye[2]);
    const InstAux& ax = aux[best.inst_slot];
    V3 P = vadd(eye, vmul(d, best.t));
    float bA = __fdiv_rn(best.U, best.det), bB = __fdiv_rn(best.V, best.det),
          bC = __fdiv_rn(best.W, best.det);
    const float* nn = ax.nrms + (size_t)best.tri_slot * 9;
    V3 no = v3(nn[0] * bA + nn[3] * bB + nn[6] * bC,
               nn[1] * bA + nn[4] * bB + nn[7] * bC,
               nn[2] * bA + nn[5] * bB + nn[8] * bC);
    const float* m = ax.nmat;
    V3 nw = v3(m[0] * no.x + m[1] * no.y + m[2] * no.z,
               m[3] * no.x + m[4] * no.y + m[5] * no.z,
               m[6] * no.x + m[7] * no.y + m[8] * no.z);
    float ln = vlen(nw);
    V3 N = ln == 0.0f ? v3(0.0f, 1.0f, 0.0f)
                      : v3(__fdiv_rn(nw.x, ln), __fdiv_rn(nw.y, ln), __fdiv_rn(nw.z, ln));
    out.position[p * 3 + 0] = P.x; out.position[p * 3 + 1] = P.y; out.position[p * 3 + 2] = P.z;
    out.normal[p * 3 + 0] = N.x; out.normal[p * 3 + 1] = N.y; out.normal[p * 3 + 2] = N.z;
    if (ax.custom_index < nmats) {
        luzrt_material mt = mats[ax.custom_index];
        out.albedo[p] = mt.albedo; out.material[p] = mt.material; out.emission[p] = mt.emission;
    } else {
        out.albedo[p] = 0xFFFFFFFFu;
        out.material[p] = 0xFFFF0080u;  // roughness 0.5, metallic 0, occlusion 1
        out.emission[p] = 0xFF000000u;
    }
}

// G-buffer texel (x, y) by a closest-hit primary ray run to completion.
template <typename STACK>
__device__ inline void primary_pixel(const Node* tlas_nodes, const InstRec* recs, const InstAux* aux,
                                     const luzrt_camera& cam, const luzrt_material* mats,
                                     uint32_t nmats, const luzrt_gbuffer_out& out, uint32_t x,
                                     uint32_t y, STACK& st, uint32_t* error_flag, bool empty) {
    Best best;
    best.found = false;
    best.t = cam.tmax;
    const V3 d = primary_dir(cam, x, y, out.width, out.height);
    if (!empty) {
        RayState r;
        ray_begin(r, st, v3(cam.eye[0], cam.eye[1], cam.eye[2]), d, cam.tmin, cam.tmax);
        for (;;) {
            int res = ray_step<true>(r, st, tlas_nodes, recs, aux, &best);
            if (res == kStepDone || (res == kStepContinue && r.next == kNone && st.sp == 0)) break;
            if (res == kStepOverflow) { atomicOr(error_flag, 1u); st.sp = st.lo = 0; break; }
        }
    }
    primary_texel(cam, d, best, aux, mats, nmats, out, x, y);
}

}  // namespace luz
