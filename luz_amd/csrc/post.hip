// post.hip — temporal accumulation (TAA) and the ACES compose of the light
// buffer on gfx950: the two passes after the lighting pass.
//
// k_taa restates taa.comp:main (:271-307) with its helpers: the 3x3 and
// plus-shaped neighbourhood min/max/mean (get_neighbor_3x3 :30-88, with the
// optional Mitchell reconstruction), the closest-depth texel of the 3x3
// neighbourhood (find_closest_3x3 :90-118), its motion vector from the
// current and previous view-projections and jitters (get_motion_vector
// :19-28), a bilinear history fetch, clip_aabb (:120-143), and the
// anti-flicker blend weights (:145-153).  Texture fetches follow the
// reference sampler (VulkanWrapper.cpp:2433-2437: linear filter, REPEAT
// addressing), so neighbourhoods wrap around the image edges.  Depth is the
// NDC depth of the G-buffer's world position through the current
// view-projection (1 for background), and world positions come from the
// position plane instead of DepthToWorld (utils.glsl:1-7) except on
// background texels, which use the far plane through the inverse
// view-projection.  Bilinear weights are exact FP32 (the hardware's are
// fixed point).
//
// k_compose restates present.frag:main for the light output (imageType 0,
// :87-90 with TonemapACES :27-35, no debug overlay): RGBA8 UNORM and,
// optionally, the FP32 tonemapped colour.
//
// One thread per pixel; every value is FP32 evaluated in source order, with
// the same order in the oracle.
#include <cmath>

#include "internal.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

struct F4 { float x, y, z, w; };

__device__ inline F4 f4(float4 v) { return F4{v.x, v.y, v.z, v.w}; }
__device__ inline F4 add4(F4 a, F4 b) { return F4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ inline F4 mul4(F4 a, float s) { return F4{a.x * s, a.y * s, a.z * s, a.w * s}; }
__device__ inline F4 div4(F4 a, float s) { return F4{a.x / s, a.y / s, a.z / s, a.w / s}; }
__device__ inline float mn(float a, float b) { return b < a ? b : a; }   // GLSL min for non-NaN
__device__ inline float mx(float a, float b) { return a < b ? b : a; }
__device__ inline F4 min4(F4 a, F4 b) { return F4{mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z), mn(a.w, b.w)}; }
__device__ inline F4 max4(F4 a, F4 b) { return F4{mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z), mx(a.w, b.w)}; }
__device__ inline bool nan4(F4 a) { return a.x != a.x || a.y != a.y || a.z != a.z || a.w != a.w; }

__device__ inline int wrap(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

struct TaaArgs {
    luzrt_taa_desc d;
};

__device__ inline F4 texel(const float* img, int x, int y, int W, int H) {
    const float4 v = reinterpret_cast<const float4*>(img)[(size_t)wrap(y, H) * W + wrap(x, W)];
    return f4(v);
}

// Column-major 4x4 (glm) times (x, y, z, w), summed left to right.
__device__ inline F4 xform4(const float* m, float x, float y, float z, float w) {
    return F4{m[0] * x + m[4] * y + m[8] * z + m[12] * w, m[1] * x + m[5] * y + m[9] * z + m[13] * w,
              m[2] * x + m[6] * y + m[10] * z + m[14] * w, m[3] * x + m[7] * y + m[11] * z + m[15] * w};
}

__device__ inline bool background(const float* normal, size_t q) {
    const float* n = normal + q * 3;
    return vlen(v3(n[0], n[1], n[2])) == 0.0f;
}

// NDC depth of texel (x, y) (wrapped); 1 on background texels.
__device__ inline float depth_at(const luzrt_taa_desc& d, int x, int y) {
    const size_t q = (size_t)wrap(y, (int)d.height) * d.width + wrap(x, (int)d.width);
    if (background(d.normal, q)) return 1.0f;
    const float* p = d.position + q * 3;
    F4 c = xform4(d.view_proj, p[0], p[1], p[2], 1.0f);
    return c.z / c.w;
}

__device__ inline float mitchell(float x) {
    const float B = 1.0f / 3.0f, C = 1.0f / 3.0f;
    float x2 = x * x;
    float x3 = x2 * x;
    return (6.0f - 2.0f * B) * x3 - (6.0f - 2.0f * B - 3.0f * C) * x2 + 1.0f;
}

__device__ inline float luminance(F4 c) { return c.x * 0.2127f + c.y * 0.7152f + c.z * 0.0722f; }

__global__ __launch_bounds__(kBlock) void k_taa(TaaArgs A) {
    const luzrt_taa_desc& d = A.d;
    const int W = (int)d.width, H = (int)d.height;
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= d.width * d.height) return;
    const int x = (int)(id % d.width), y = (int)(id / d.width);
    const float sw_ = (float)W, sh_ = (float)H;
    const float su = ((float)x + 0.5f) / sw_, sv = ((float)y + 0.5f) / sh_;

    // find_closest_3x3: the first strictly nearer texel in row order
    int bx = -1, by = -1;
    float bz = depth_at(d, x - 1, y - 1);
    for (int k = 1; k < 9; ++k) {
        int ox = k % 3 - 1, oy = k / 3 - 1;
        float z = depth_at(d, x + ox, y + oy);
        if (bz > z) { bz = z; bx = ox; by = oy; }
    }
    const float ddx = fabsf(1.0f / sw_), ddy = fabsf(1.0f / sh_);
    const float cu = su + ddx * (float)bx, cv = sv + ddy * (float)by;

    // get_motion_vector at the closest texel
    float wx, wy, wz;
    {
        const size_t q = (size_t)wrap(y + by, H) * d.width + wrap(x + bx, W);
        if (!background(d.normal, q)) {
            const float* p = d.position + q * 3;
            wx = p[0]; wy = p[1]; wz = p[2];
        } else {
            F4 w = xform4(d.inv_view_proj, cu * 2.0f - 1.0f, cv * 2.0f - 1.0f, 1.0f, 1.0f);
            wx = w.x / w.w; wy = w.y / w.w; wz = w.z / w.w;
        }
    }
    F4 pn = xform4(d.prev_view_proj, wx, wy, wz, 1.0f);
    F4 cn = xform4(d.view_proj, wx, wy, wz, 1.0f);
    float pnx = pn.x / pn.w, pny = pn.y / pn.w;
    float cnx = cn.x / cn.w, cny = cn.y / cn.w;
    float mvx = ((cnx - d.jitter[0]) - (pnx - d.prev_jitter[0])) * 0.5f;
    float mvy = ((cny - d.jitter[1]) - (pny - d.prev_jitter[1])) * 0.5f;
    const float hu = su - mvx, hv = sv - mvy;

    // get_neighbor_3x3
    F4 c[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) c[k] = texel(d.light, x + k % 3 - 1, y + k / 3 - 1, W, H);
    F4 cmin = min4(c[0], min4(c[1], min4(c[2], min4(c[3], min4(c[4], min4(c[5], min4(c[6], min4(c[7], c[8]))))))));
    F4 cmax = max4(c[0], max4(c[1], max4(c[2], max4(c[3], max4(c[4], max4(c[5], max4(c[6], max4(c[7], c[8]))))))));
    F4 cavg = c[0];
#pragma unroll
    for (int k = 1; k < 9; ++k) cavg = add4(cavg, c[k]);
    cavg = div4(cavg, 9.0f);
    F4 cmin5 = min4(c[1], min4(c[3], min4(c[4], min4(c[5], c[7]))));
    F4 cmax5 = max4(c[1], max4(c[3], max4(c[4], max4(c[5], c[7]))));
    F4 cavg5 = div4(add4(add4(add4(add4(c[1], c[3]), c[4]), c[5]), c[7]), 5.0f);
    cmin = mul4(add4(cmin, cmin5), 0.5f);
    cmax = mul4(add4(cmax, cmax5), 0.5f);
    cavg = mul4(add4(cavg, cavg5), 0.5f);
    F4 src = F4{0.0f, 0.0f, 0.0f, 0.0f};
    if (d.reconstruct == 1) {
        float wsum = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            float ox = (float)(k % 3 - 1), oy = (float)(k / 3 - 1);
            float w = mitchell(sqrtf(ox * ox + oy * oy));
            src = add4(src, mul4(c[k], w));
            wsum = wsum + w;
        }
        src = div4(src, wsum);
    }
    if (d.reconstruct == 0 || nan4(src)) src = c[4];

    float4* out = reinterpret_cast<float4*>(d.out);
    if (hu > 1.0f || hv > 1.0f || hu < 0.0f || hv < 0.0f) {
        out[id] = make_float4(src.x, src.y, src.z, src.w);
        return;
    }
    // bilinear history fetch (REPEAT addressing)
    float tx = hu * sw_ - 0.5f, ty = hv * sh_ - 0.5f;
    float fx0 = floorf(tx), fy0 = floorf(ty);
    float ax = tx - fx0, ay = ty - fy0;
    int ix = (int)fx0, iy = (int)fy0;
    F4 h00 = texel(d.history, ix, iy, W, H), h10 = texel(d.history, ix + 1, iy, W, H);
    F4 h01 = texel(d.history, ix, iy + 1, W, H), h11 = texel(d.history, ix + 1, iy + 1, W, H);
    F4 top = add4(mul4(h00, 1.0f - ax), mul4(h10, ax));
    F4 bot = add4(mul4(h01, 1.0f - ax), mul4(h11, ax));
    F4 hist = add4(mul4(top, 1.0f - ay), mul4(bot, ay));

    // clip_aabb(cmin.rgb, cmax.rgb, clamp(cavg, cmin, cmax), history)
    F4 p = min4(max4(cavg, cmin), cmax);
    F4 r = F4{hist.x - p.x, hist.y - p.y, hist.z - p.z, hist.w - p.w};
    const float rmaxx = cmax.x - p.x, rmaxy = cmax.y - p.y, rmaxz = cmax.z - p.z;
    const float rminx = cmin.x - p.x, rminy = cmin.y - p.y, rminz = cmin.z - p.z;
    const float eps = 0.00000001f;
    if (r.x > rmaxx + eps) r = mul4(r, rmaxx / r.x);
    if (r.y > rmaxy + eps) r = mul4(r, rmaxy / r.y);
    if (r.z > rmaxz + eps) r = mul4(r, rmaxz / r.z);
    if (r.x < rminx - eps) r = mul4(r, rminx / r.x);
    if (r.y < rminy - eps) r = mul4(r, rminy / r.y);
    if (r.z < rminz - eps) r = mul4(r, rminz / r.z);
    hist = add4(p, r);

    float sw = 0.05f;
    float hw = 1.0f - sw;
    {
        float ms = mx(mx(src.x, src.y), src.z) + 1.0f, mh = mx(mx(hist.x, hist.y), hist.z) + 1.0f;
        float rs = 1.0f / ms, rh = 1.0f / mh;
        float ls = luminance(F4{src.x * rs, src.y * rs, src.z * rs, 0.0f});
        float lh = luminance(F4{hist.x * rh, hist.y * rh, hist.z * rh, 0.0f});
        sw = sw * (1.0f / (1.0f + ls));
        hw = hw * (1.0f / (1.0f + lh));
    }
    F4 res = div4(add4(mul4(src, sw), mul4(hist, hw)), mx(sw + hw, 0.0000001f));
    if (nan4(res)) res = src;
    out[id] = make_float4(res.x, res.y, res.z, res.w);
}

__device__ inline float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float v = (x * (a * x + b)) / (x * (c * x + d) + e);
    return powf(v, 1.0f / 2.2f);
}

__device__ inline uint32_t unorm8(float v) {
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);   // NaN -> 0
    return (uint32_t)(v * 255.0f + 0.5f);
}

__global__ __launch_bounds__(kBlock) void k_compose(const float4* light, uint32_t npix, uint32_t* rgba8,
                                                    float4* rgb) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float4 v = light[p];
    float r = aces(v.x), g = aces(v.y), b = aces(v.z);
    if (rgba8) rgba8[p] = unorm8(r) | unorm8(g) << 8 | unorm8(b) << 16 | 255u << 24;
    if (rgb) rgb[p] = make_float4(r, g, b, 1.0f);
}

}  // namespace
}  // namespace luz

using namespace luz;

extern "C" luzrt_status luzrt_taa(luzrt_ctx* ctx, const luzrt_taa_desc* d) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !d->light || !d->history || !d->out || !d->position || !d->normal)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_taa: null argument");
    if (d->width == 0 || d->height == 0 || d->width > 16384 || d->height > 16384)
        return fail(ctx, LUZRT_E_LIMIT, "luzrt_taa: size %ux%u outside 1..16384", d->width, d->height);
    if (((uintptr_t)d->light | (uintptr_t)d->history | (uintptr_t)d->out) & 15u)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_taa: RGBA32F images must be 16-byte aligned");
    if (d->out == d->light || d->out == d->history)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_taa: the output must not alias an input");
    const uint32_t npix = d->width * d->height;
    TaaArgs A{*d};
    stage_begin(ctx, kStagePost);
    k_taa<<<(npix + kBlock - 1) / kBlock, kBlock, 0, ctx->stream>>>(A);
    stage_end(ctx, kStagePost);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

extern "C" luzrt_status luzrt_compose(luzrt_ctx* ctx, uint32_t width, uint32_t height, const float* light,
                                      uint32_t* rgba8_out, float* rgba32f_out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!light || (!rgba8_out && !rgba32f_out))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_compose: null argument");
    if (((uintptr_t)light | (uintptr_t)rgba32f_out) & 15u)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_compose: RGBA32F images must be 16-byte aligned");
    const uint64_t npix = (uint64_t)width * height;
    if (npix == 0 || npix > 0xFFFFFFFFull) return fail(ctx, LUZRT_E_LIMIT, "luzrt_compose: bad size");
    stage_begin(ctx, kStagePost);
    k_compose<<<(uint32_t)((npix + kBlock - 1) / kBlock), kBlock, 0, ctx->stream>>>(
        (const float4*)light, (uint32_t)npix, rgba8_out, (float4*)rgba32f_out);
    stage_end(ctx, kStagePost);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}
