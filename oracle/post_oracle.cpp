// post_oracle.cpp — CPU oracle (test infrastructure only) of the passes after
// the lighting pass: temporal accumulation (taa.comp:main :271-307 and its
// helpers :19-153) and the ACES compose of the light output
// (present.frag:27-35, 87-90).  FP32 evaluated in the written order
// (-ffp-contract=off), images addressed with REPEAT wrapping as the
// reference sampler (VulkanWrapper.cpp:2433-2437), depth = NDC depth of the
// position plane (1 for background), background world positions on the far
// plane through the inverse view-projection, exact FP32 bilinear weights.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "luz_oracle.h"

namespace {

struct C4 {
    float v[4];
};

inline int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

inline C4 fetch(const float* img, int x, int y, int W, int H) {
    C4 c;
    std::memcpy(c.v, img + ((size_t)wrapi(y, H) * W + wrapi(x, W)) * 4, 16);
    return c;
}

// GLSL min/max on non-NaN values
inline float gmin(float a, float b) { return b < a ? b : a; }
inline float gmax(float a, float b) { return a < b ? b : a; }

inline void mat_vec(const float* m, const float in[4], float out[4]) {
    for (int r = 0; r < 4; ++r) {
        float acc = m[r] * in[0];
        acc = acc + m[4 + r] * in[1];
        acc = acc + m[8 + r] * in[2];
        acc = acc + m[12 + r] * in[3];
        out[r] = acc;
    }
}

inline bool is_background(const float* normal, size_t q) {
    const float* n = normal + q * 3;
    float s = n[0] * n[0] + n[1] * n[1];
    s = s + n[2] * n[2];
    return std::sqrt(s) == 0.0f;
}

float ndc_depth(const orc_taa_desc* d, int x, int y) {
    const size_t q = (size_t)wrapi(y, (int)d->height) * d->width + wrapi(x, (int)d->width);
    if (is_background(d->normal, q)) return 1.0f;
    const float in[4] = {d->position[q * 3], d->position[q * 3 + 1], d->position[q * 3 + 2], 1.0f};
    float c[4];
    mat_vec(d->view_proj, in, c);
    return c[2] / c[3];
}

float mitchell_filter(float x) {   // utils.glsl MitchellFilter, B = C = 1/3
    const float B = 1.0f / 3.0f, Cc = 1.0f / 3.0f;
    float x2 = x * x;
    float x3 = x2 * x;
    return (6.0f - 2.0f * B) * x3 - (6.0f - 2.0f * B - 3.0f * Cc) * x2 + 1.0f;
}

float luma(const float c[3]) {   // utils.glsl Luminance: dot with (0.2127, 0.7152, 0.0722)
    float l = c[0] * 0.2127f + c[1] * 0.7152f;
    return l + c[2] * 0.0722f;
}

bool has_nan(const C4& c) {
    return std::isnan(c.v[0]) || std::isnan(c.v[1]) || std::isnan(c.v[2]) || std::isnan(c.v[3]);
}

void taa_pixel(const orc_taa_desc* d, int x, int y, float* out) {
    const int W = (int)d->width, H = (int)d->height;
    const float fw = (float)W, fh = (float)H;
    const float su = ((float)x + 0.5f) / fw, sv = ((float)y + 0.5f) / fh;

    // find_closest_3x3 (taa.comp:90-118)
    const int order[9][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {0, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
    int best = 0;
    float bestz = ndc_depth(d, x - 1, y - 1);
    for (int k = 1; k < 9; ++k) {
        float z = ndc_depth(d, x + order[k][0], y + order[k][1]);
        if (bestz > z) {
            bestz = z;
            best = k;
        }
    }
    const float dx = std::fabs(1.0f / fw), dy = std::fabs(1.0f / fh);
    const float cu = su + dx * (float)order[best][0], cv = sv + dy * (float)order[best][1];

    // get_motion_vector (taa.comp:19-28)
    float wp[3];
    const size_t q = (size_t)wrapi(y + order[best][1], H) * d->width + wrapi(x + order[best][0], W);
    if (!is_background(d->normal, q)) {
        for (int k = 0; k < 3; ++k) wp[k] = d->position[q * 3 + k];
    } else {
        const float clip[4] = {cu * 2.0f - 1.0f, cv * 2.0f - 1.0f, 1.0f, 1.0f};
        float w[4];
        mat_vec(d->inv_view_proj, clip, w);
        for (int k = 0; k < 3; ++k) wp[k] = w[k] / w[3];
    }
    const float p4[4] = {wp[0], wp[1], wp[2], 1.0f};
    float prev[4], cur[4];
    mat_vec(d->prev_view_proj, p4, prev);
    mat_vec(d->view_proj, p4, cur);
    const float mv[2] = {((cur[0] / cur[3] - d->jitter[0]) - (prev[0] / prev[3] - d->prev_jitter[0])) * 0.5f,
                         ((cur[1] / cur[3] - d->jitter[1]) - (prev[1] / prev[3] - d->prev_jitter[1])) * 0.5f};
    const float hu = su - mv[0], hv = sv - mv[1];

    // get_neighbor_3x3 (taa.comp:30-88)
    C4 n[9];
    for (int k = 0; k < 9; ++k) n[k] = fetch(d->light, x + order[k][0], y + order[k][1], W, H);
    float cmin[4], cmax[4], cavg[4], src[4];
    for (int c = 0; c < 4; ++c) {
        // min(ctl, min(ctc, ... min(cbc, cbr))): innermost first
        float mn = n[8].v[c], mx = n[8].v[c];
        for (int k = 7; k >= 0; --k) {
            mn = gmin(n[k].v[c], mn);
            mx = gmax(n[k].v[c], mx);
        }
        float sum = n[0].v[c];
        for (int k = 1; k < 9; ++k) sum = sum + n[k].v[c];
        float avg = sum / 9.0f;
        const int plus[5] = {1, 3, 4, 5, 7};
        float mn5 = n[7].v[c], mx5 = n[7].v[c];
        for (int k = 3; k >= 0; --k) {
            mn5 = gmin(n[plus[k]].v[c], mn5);
            mx5 = gmax(n[plus[k]].v[c], mx5);
        }
        float s5 = n[1].v[c];
        for (int k = 1; k < 5; ++k) s5 = s5 + n[plus[k]].v[c];
        float avg5 = s5 / 5.0f;
        cmin[c] = 0.5f * (mn + mn5);
        cmax[c] = 0.5f * (mx + mx5);
        cavg[c] = 0.5f * (avg + avg5);
    }
    C4 s{{0.0f, 0.0f, 0.0f, 0.0f}};
    if (d->reconstruct == 1) {
        float wsum = 0.0f;
        for (int k = 0; k < 9; ++k) {
            float ox = (float)order[k][0], oy = (float)order[k][1];
            float w = mitchell_filter(std::sqrt(ox * ox + oy * oy));
            for (int c = 0; c < 4; ++c) s.v[c] = s.v[c] + n[k].v[c] * w;
            wsum = wsum + w;
        }
        for (int c = 0; c < 4; ++c) s.v[c] = s.v[c] / wsum;
    }
    if (d->reconstruct == 0 || has_nan(s)) s = n[4];
    for (int c = 0; c < 4; ++c) src[c] = s.v[c];

    if (hu > 1.0f || hv > 1.0f || hu < 0.0f || hv < 0.0f) {
        std::memcpy(out, src, 16);
        return;
    }
    // bilinear history sample
    const float tx = hu * fw - 0.5f, ty = hv * fh - 0.5f;
    const float x0 = std::floor(tx), y0 = std::floor(ty);
    const float ax = tx - x0, ay = ty - y0;
    const int ix = (int)x0, iy = (int)y0;
    C4 h00 = fetch(d->history, ix, iy, W, H), h10 = fetch(d->history, ix + 1, iy, W, H);
    C4 h01 = fetch(d->history, ix, iy + 1, W, H), h11 = fetch(d->history, ix + 1, iy + 1, W, H);
    float hist[4];
    for (int c = 0; c < 4; ++c) {
        float top = h00.v[c] * (1.0f - ax) + h10.v[c] * ax;
        float bot = h01.v[c] * (1.0f - ax) + h11.v[c] * ax;
        hist[c] = top * (1.0f - ay) + bot * ay;
    }
    // clip_aabb(cmin.rgb, cmax.rgb, clamp(cavg, cmin, cmax), history) (taa.comp:120-143)
    float p[4], r[4];
    for (int c = 0; c < 4; ++c) {
        p[c] = gmin(gmax(cavg[c], cmin[c]), cmax[c]);
        r[c] = hist[c] - p[c];
    }
    const float eps = 0.00000001f;
    auto scale = [&](float f) { for (int c = 0; c < 4; ++c) r[c] = r[c] * f; };
    for (int c = 0; c < 3; ++c) {
        float rmax = cmax[c] - p[c];
        if (r[c] > rmax + eps) scale(rmax / r[c]);
    }
    for (int c = 0; c < 3; ++c) {
        float rmin = cmin[c] - p[c];
        if (r[c] < rmin - eps) scale(rmin / r[c]);
    }
    for (int c = 0; c < 4; ++c) hist[c] = p[c] + r[c];

    // anti_flicker weights (taa.comp:145-153) and the blend (:296-305)
    float sw = 0.05f;
    float hw = 1.0f - sw;
    float cs[3], ch[3];
    float ks = 1.0f / (gmax(gmax(src[0], src[1]), src[2]) + 1.0f);
    float kh = 1.0f / (gmax(gmax(hist[0], hist[1]), hist[2]) + 1.0f);
    for (int c = 0; c < 3; ++c) {
        cs[c] = src[c] * ks;
        ch[c] = hist[c] * kh;
    }
    sw = sw * (1.0f / (1.0f + luma(cs)));
    hw = hw * (1.0f / (1.0f + luma(ch)));
    float den = gmax(sw + hw, 0.0000001f);
    float res[4];
    bool bad = false;
    for (int c = 0; c < 4; ++c) {
        res[c] = (src[c] * sw + hist[c] * hw) / den;
        bad = bad || std::isnan(res[c]);
    }
    std::memcpy(out, bad ? src : res, 16);
}

float aces_gamma(float x) {   // present.frag TonemapACES
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float v = (x * (a * x + b)) / (x * (c * x + d) + e);
    return std::pow(v, 1.0f / 2.2f);
}

uint32_t to_unorm8(float v) {
    if (!(v > 0.0f)) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    return (uint32_t)(v * 255.0f + 0.5f);
}

}  // namespace

extern "C" int orc_taa(const orc_taa_desc* d, float* out) {
    if (!d || !out || !d->light || !d->history || !d->position || !d->normal) return -1;
    for (uint32_t y = 0; y < d->height; ++y)
        for (uint32_t x = 0; x < d->width; ++x)
            taa_pixel(d, (int)x, (int)y, out + ((size_t)y * d->width + x) * 4);
    return 0;
}

extern "C" int orc_compose(const float* light, uint32_t npix, uint32_t* rgba8, float* rgb) {
    if (!light) return -1;
    for (uint32_t p = 0; p < npix; ++p) {
        float r = aces_gamma(light[p * 4]), g = aces_gamma(light[p * 4 + 1]), b = aces_gamma(light[p * 4 + 2]);
        if (rgba8) rgba8[p] = to_unorm8(r) | to_unorm8(g) << 8 | to_unorm8(b) << 16 | 255u << 24;
        if (rgb) {
            rgb[p * 4] = r;
            rgb[p * 4 + 1] = g;
            rgb[p * 4 + 2] = b;
            rgb[p * 4 + 3] = 1.0f;
        }
    }
    return 0;
}
