// denoise.hip — learned per-pixel denoiser of 1-spp ambient occlusion and
// soft shadows on gfx950, inference and training with BF16 MFMA.
//
// Network (per (pixel, light) sample): 150 features = a 5x5 neighbourhood
// (clamped to the image) x 6 channels {AO, visibility of the light, relative
// depth, normal.xyz}; 4 hidden layers of 64 with ReLU; 2 outputs {AO,
// visibility} added to the 5x5 mean of the noisy AO and visibility inputs over
// the window's foreground pixels (FP32), so the network learns a correction to
// that box filter and BF16 precision bounds only the correction.  A pixel's AO
// is the mean of the AO outputs of its L samples.  Weights are BF16,
// accumulation FP32, biases FP32.
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, 64 lanes): weights are the A
// operand (16 output features x 32 inputs), samples the B operand (32 inputs
// x 16 samples).  The 16x16 FP32 result leaves lane (j, g) = (lane % 16,
// lane / 16) holding features 4g..4g+3 of sample j, which is exactly the B
// operand lane layout of the next layer when its input index k of element e
// of K-step s is  k = 16 * (2s + e/4) + 4g + e%4.  Weight fragments are
// packed in that order, so activations stay in registers through all five
// layers, forward and backward (the transposed weights W^T are packed the
// same way for the backward pass).  The features of layer 1 use k = 32s +
// 8g + e.  Any K order works as long as A and B agree, because the hardware
// pairs element e of lane group g of A with the same element of B.
//
// Inference (k_dn_infer) takes 16-pixel tiles: the 125 features that do not
// depend on the light enter layer 1 once per pixel, the 25 visibilities of a
// light as one more K-step per light, and layers 2..5 run per light.
//
// Training: k_dn_forward runs forward, loss and the backward pass to
// every layer's pre-activation gradient dz, and stores x, h and dz (BF16)
// sample-minor ([feature][sample]); k_dn_wgrad forms dW = dz^T h per chunk of
// 1024 samples with MFMA (K = samples) and bias sums; k_dn_reduce adds the
// chunk partials in a fixed order (deterministic gradients); k_dn_adam and
// k_dn_pack update the FP32 master weights and re-pack the BF16 fragments.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace luz {

struct DnFrameDev {
    uint32_t width, height, num_lights, pad;
    const float* position;
    const float* normal;
    const uint8_t* shadow;
    const uint8_t* ao;
    float cam[3];
    float shadow_n, ao_n;
    float pad2;
    // per-pixel planes of the prepass: (camera distance or -1 on background,
    // normal.xyz) and the AO fraction
    const float4* pix;
    const float* ao_frac;
};

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = LUZRT_DN_FEATURES, kHid = LUZRT_DN_HIDDEN, kOut = 2;
constexpr int kBlock = 256;
constexpr uint32_t kChunkRows = 1024;
// inference: one 16-pixel tile per wave at a time, 8 waves per group sharing
// one copy of the weights in LDS
constexpr int kInferBlock = 512;
constexpr uint32_t kInferRounds = 4;   // bands of tiles per chunk an XCD sweeps in order

constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn;
constexpr int oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid;
constexpr int oW3 = oB2 + kHid, oB3 = oW3 + kHid * kHid;
constexpr int oW4 = oB3 + kHid, oB4 = oW4 + kHid * kHid;
constexpr int oW5 = oB4 + kHid, oB5 = oW5 + kOut * kHid;
constexpr int kParams = oB5 + kOut;
static_assert(kParams == LUZRT_DN_PARAMS, "parameter count");
constexpr int kWOff[5] = {oW1, oW2, oW3, oW4, oW5};
constexpr int kBOff[5] = {oB1, oB2, oB3, oB4, oB5};

// Fragment table (one fragment = 64 lanes x 8 BF16 = 1 KB):
//   F1(t,s) t<4 s<5 | F2..F4(t,s) t<4 s<2 | G2..G4(t,s) | G5(t) t<4
//   | FG(t,s) t<4 s<4 | FV(t) t<4
// Layer 5 (64 -> 2) runs in FP32 (dn_out_fp32); G5 holds W5^T for the backward pass.
// FG and FV hold layer 1 again in the inference order, which splits the
// features of a (pixel, light) sample into the 125 that do not depend on the
// light (one product per pixel) and the 25 visibilities (one K-step per light).
// Lane group g only ever reads neighbours 8g .. 8g + 7 (those below 25):
//   FG, K-steps s < 4: element e is channel (AO, n.x, n.y, n.z)[e % 4] of
//       neighbour 8g + 2s + e / 4;
//   FG, K-step 4: element e is the relative depth of neighbour 8g + e;
//   FV: element e is the visibility of neighbour 8g + e.
// F5(s) s < 2 holds layer 5 for inference, where it runs on MFMA: output row 0
// and 1 are the BF16 rounding of W5's rows, rows 2 and 3 the BF16 rounding of
// what that leaves (W5 - hi), rows 4..15 zero, K in the hidden-layer order.
constexpr int kF1 = 0, kF2 = 20, kF3 = 28, kF4 = 36, kG2 = 44, kG3 = 52, kG4 = 60,
              kG5 = 68, kFG = 72, kFV = 92, kF5 = 96, kFragTrain = 72, kFragAll = 98;
// inference fragments in LDS: F2..F4 (24), then FG and FV (24), then F5 (2)
constexpr int kIF2 = 0, kIF3 = 8, kIF4 = 16, kIFG = 24, kIFV = 44, kIF5 = 48, kFragInfer = 50;

// Original feature index (6 n + channel) of element e of lane group g in
// inference K-step s (s == 5: the visibility step), or -1 for padding.
__host__ __device__ inline int infer_feature(int s, int g, int e) {
    if (s < 4) {
        const int n = 8 * g + 2 * s + (e >> 2), ch = e & 3;
        return n < 25 ? 6 * n + (ch == 0 ? 0 : ch + 2) : -1;
    }
    const int n = 8 * g + e;
    if (n >= 25) return -1;
    return s == 4 ? 6 * n + 2 : 6 * n + 1;
}

__host__ __device__ inline int khid(int s, int g, int e) { return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// BF16 of two finite floats, round to nearest even (v_cvt_pk_bf16_f32): a in
// the low half, b in the high half.
__device__ inline uint32_t bf16_pack2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{a, b}), bf16x2));
}
__device__ inline uint32_t bf16_bits(float f) { return bf16_pack2(f, 0.0f) & 0xFFFFu; }
// max(z, 0) for finite z on the bit pattern (negative floats are negative integers)
__device__ inline float relu_bits(float z) {
    const int b = __float_as_int(z);
    return __int_as_float(b > 0 ? b : 0);
}
__device__ inline float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }
__device__ inline float clamp01f(float a) { return a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a); }
__device__ inline bf16x8 as_bf16x8(uint4 q) { return __builtin_bit_cast(bf16x8, q); }
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct DnArgs {
    DnFrameDev F;
    const uint4* frags;
    const float* params;
    uint32_t nsamples;
    uint32_t lights;           // max(L, 1)
    const uint32_t* samples;   // pixel | light << 24
    const float* target_ao;    // [npix]
    const float* target_vis;   // [L][npix]
    uint32_t row_base, ldn;    // training rows of this call, leading dimension
    uint16_t* XT;              // [160][ldn]
    uint16_t* HT;              // [4][64][ldn]
    uint16_t* ZT;              // [4][64][ldn]
    uint16_t* YT;              // [16][ldn]
    float grad_scale;          // dLoss/dy = grad_scale * (y - target)
    float* loss_sum;
};

// The 6 channels of neighbour (qx, qy) of a pixel at distance dc from the camera.
struct Nb { float c0, c1, c2, c3, c4, c5; };

__device__ inline Nb dn_neighbour(const DnFrameDev& F, uint32_t qx, uint32_t qy, uint32_t light, float dc) {
    const size_t q = (size_t)qy * F.width + qx;
    const float4 px = F.pix[q];
    Nb r;
    if (px.x < 0.0f) {   // background
        r.c0 = 0.0f; r.c1 = 0.0f; r.c2 = 1.0f; r.c3 = 0.0f; r.c4 = 0.0f; r.c5 = 0.0f;
        return r;
    }
    const float d = px.x;
    r.c0 = F.ao_frac[q];
    r.c1 = F.num_lights ? 1.0f - (float)F.shadow[(size_t)light * F.width * F.height + q] / F.shadow_n
                        : 0.0f;
    float rel = dc > 0.0f ? (d - dc) / dc * 16.0f : 0.0f;
    r.c2 = rel < -1.0f ? -1.0f : (rel > 1.0f ? 1.0f : rel);
    r.c3 = px.y; r.c4 = px.z; r.c5 = px.w;
    return r;
}

// Prepass: camera distance (-1 on background pixels, whose normal is zero),
// normal and AO fraction of every pixel, each computed once per frame.
__global__ __launch_bounds__(256) void k_dn_prep(DnFrameDev F, float4* pix, float* ao_frac) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= F.width * F.height) return;
    const float* n = F.normal + (size_t)p * 3;
    const float nx = n[0], ny = n[1], nz = n[2];
    if (vlen(v3(nx, ny, nz)) == 0.0f) {
        pix[p] = make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
        ao_frac[p] = 0.0f;
        return;
    }
    const float* P = F.position + (size_t)p * 3;
    const float d = vlen(v3(P[0] - F.cam[0], P[1] - F.cam[1], P[2] - F.cam[2]));
    pix[p] = make_float4(d, nx, ny, nz);
    ao_frac[p] = (float)F.ao[p] / F.ao_n;
}

// The base the network's outputs are added to: the mean, over the foreground
// pixels of the clamped 5x5 window, of the AO fraction and of the visibility
// of `light` (0 without lights).  Summed per group of 8 neighbours (n = 8g + e,
// in order of e), then ((g0 + g1) + (g2 + g3)): the order of the lane
// reduction in k_dn_infer, so training and inference bases are identical.
__device__ inline void dn_box(const DnFrameDev& F, uint32_t x, uint32_t y, uint32_t light, float& bao,
                              float& bvis) {
    float sa[4], sv[4];
    int cnt = 0;
    for (int g = 0; g < 4; ++g) {
        sa[g] = 0.0f;
        sv[g] = 0.0f;
        for (int e = 0; e < 8; ++e) {
            const int n = 8 * g + e;
            if (n >= 25) break;
            int qx = (int)x + n % 5 - 2, qy = (int)y + n / 5 - 2;
            qx = qx < 0 ? 0 : (qx >= (int)F.width ? (int)F.width - 1 : qx);
            qy = qy < 0 ? 0 : (qy >= (int)F.height ? (int)F.height - 1 : qy);
            const size_t q = (size_t)qy * F.width + qx;
            if (F.pix[q].x < 0.0f) continue;
            cnt++;
            sa[g] += F.ao_frac[q];
            if (F.num_lights)
                sv[g] += 1.0f - (float)F.shadow[(size_t)light * F.width * F.height + q] / F.shadow_n;
        }
    }
    const float c = (float)(cnt ? cnt : 1);
    bao = ((sa[0] + sa[1]) + (sa[2] + sa[3])) / c;
    bvis = ((sv[0] + sv[1]) + (sv[2] + sv[3])) / c;
}

__device__ inline float nb_channel(const Nb& b, int c) {
    return c == 0 ? b.c0 : c == 1 ? b.c1 : c == 2 ? b.c2 : c == 3 ? b.c3 : c == 4 ? b.c4 : b.c5;
}

// Features f0 .. f0+7 (BF16 bits, element e in half e of the uint4) of the sample
// at pixel (x, y), light `light`; zero beyond 150 and for invalid samples.
__device__ inline uint4 dn_features(const DnFrameDev& F, bool valid, uint32_t x, uint32_t y,
                                   uint32_t light, float dc, int f0) {
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    if (valid) {
        int cur = -1;
        Nb nb{};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            int f = f0 + e;
            if (f < kIn) {
                int n = f / 6, c = f - n * 6;
                if (n != cur) {
                    int dx = n % 5 - 2, dy = n / 5 - 2;
                    int qx = (int)x + dx, qy = (int)y + dy;
                    qx = qx < 0 ? 0 : (qx >= (int)F.width ? (int)F.width - 1 : qx);
                    qy = qy < 0 ? 0 : (qy >= (int)F.height ? (int)F.height - 1 : qy);
                    nb = dn_neighbour(F, (uint32_t)qx, (uint32_t)qy, light, dc);
                    cur = n;
                }
                w[e >> 1] |= bf16_bits(nb_channel(nb, c)) << (16 * (e & 1));
            }
        }
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// Bias + ReLU of a hidden layer's 4x4 tiles of one sample column (training);
// the next layer's B fragments, the ReLU mask (bit t*4+i) and h stored.
__device__ inline void dn_hidden_epilogue(const f32x4 (&acc)[4], const float* bias, int g, uint4 (&hb)[2],
                                          uint32_t& mask, uint16_t* HTl, uint32_t ldn, uint32_t row,
                                          bool store) {
    uint32_t w[2][4];
    mask = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float z = acc[t][i] + bias[16 * t + 4 * g + i];
            const bool on = z > 0.0f;
            h[i] = on ? z : 0.0f;
            mask |= (on ? 1u : 0u) << (t * 4 + i);
        }
        // elements e = (t & 1) * 4 + i of K-step s = t / 2
#pragma unroll
        for (int i = 0; i < 4; i += 2) {
            const uint32_t bits = bf16_pack2(h[i], h[i + 1]);
            w[t >> 1][((t & 1) * 4 + i) >> 1] = bits;
            if (store) {
                const int o = 16 * t + 4 * g + i;
                HTl[(size_t)o * ldn + row] = (uint16_t)bits;
                HTl[(size_t)(o + 1) * ldn + row] = (uint16_t)(bits >> 16);
            }
        }
    }
    hb[0] = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
    hb[1] = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
}

typedef short s16x2 __attribute__((ext_vector_type(2)));

// ReLU of a hidden layer's 4x4 tiles (bias already in the accumulators) into
// the next layer's B fragments (inference).  The ReLU is taken after rounding,
// on each BF16 half as a signed 16-bit integer (v_pk_max_i16 with 0): a value
// with its sign bit set, -0 and negative NaN included, becomes +0, as
// relu_bits before the rounding gives, and every other value is kept.
__device__ inline void dn_relu_pack(const f32x4 (&acc)[4], uint4 (&hb)[2]) {
    uint32_t w[2][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; i += 2) {
            const s16x2 h = __builtin_bit_cast(s16x2, bf16_pack2(acc[t][i], acc[t][i + 1]));
            w[t >> 1][((t & 1) * 4 + i) >> 1] =
                __builtin_bit_cast(uint32_t, __builtin_elementwise_max(h, (s16x2){0, 0}));
        }
    hb[0] = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
    hb[1] = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
}

// Layer 5 in FP32: the two outputs (without bias) of sample column j from the
// FP32 activations of layer 4 (lane (j, g) holds features 16t + 4g + i) and
// FP32 weights w5 ([2][64]); per lane in (t, i) order, then over the lane
// groups as ((g0 + g1) + (g2 + g3)).  FP32 keeps a correction that the weights
// cancel to about zero as precise as the box base it is added to.
__device__ inline void dn_out_fp32(const f32x4 (&h)[4], const float* w5, int g, float& y0, float& y1) {
    float p0 = 0.0f, p1 = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 a = *reinterpret_cast<const f32x4*>(&w5[16 * t + 4 * g]);
        const f32x4 b = *reinterpret_cast<const f32x4*>(&w5[kHid + 16 * t + 4 * g]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            p0 = fmaf(a[i], h[t][i], p0);
            p1 = fmaf(b[i], h[t][i], p1);
        }
    }
    p0 += __shfl_xor(p0, 16);
    p1 += __shfl_xor(p1, 16);
    p0 += __shfl_xor(p0, 32);
    p1 += __shfl_xor(p1, 32);
    y0 = p0;
    y1 = p1;
}

// Layer (k >= 2) with 2 K-steps: acc[q][t] = sum_s frag(t, s) * hb[q][s].
template <int Q>
__device__ inline void dn_layer64(const uint4* sfrag, int fbase, int lane, const uint4 (&hb)[Q][2],
                                  f32x4 (&acc)[Q][4]) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[q][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            bf16x8 a = as_bf16x8(sfrag[(fbase + t * 2 + s) * 64 + lane]);
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q][t] = mfma(a, as_bf16x8(hb[q][s]), acc[q][t]);
        }
}

// Backward through a hidden layer's ReLU: dz = dh * mask; stores dz (BF16)
// when training rows are valid, and builds the B fragments of dz.
__device__ inline void dn_relu_back(const f32x4 (&dh)[4], uint32_t mask, int g, uint4 (&zb)[2],
                                    uint16_t* ZTl, uint32_t ldn, uint32_t row, bool store) {
    uint32_t w[2][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; i += 2) {
            const float d0 = (mask >> (t * 4 + i)) & 1u ? dh[t][i] : 0.0f;
            const float d1 = (mask >> (t * 4 + i + 1)) & 1u ? dh[t][i + 1] : 0.0f;
            const uint32_t bits = bf16_pack2(d0, d1);
            w[t >> 1][((t & 1) * 4 + i) >> 1] = bits;
            if (store) {
                const int o = 16 * t + 4 * g + i;
                ZTl[(size_t)o * ldn + row] = (uint16_t)bits;
                ZTl[(size_t)(o + 1) * ldn + row] = (uint16_t)(bits >> 16);
            }
        }
    zb[0] = make_uint4(w[0][0], w[0][1], w[0][2], w[0][3]);
    zb[1] = make_uint4(w[1][0], w[1][1], w[1][2], w[1][3]);
}

// Training forward, loss and backward of explicit samples.  Q: 16-sample
// tiles per wave and iteration (each weight fragment read from LDS feeds Q
// MFMAs).
template <int Q>
__global__ __launch_bounds__(kBlock) void k_dn_forward(DnArgs A) {
    __shared__ uint4 sfrag[kFragTrain * 64];
    __shared__ float sbias[4 * kHid + kOut];
    __shared__ __align__(16) float sw5[kOut * kHid];
    for (int i = threadIdx.x; i < kFragTrain * 64; i += kBlock) sfrag[i] = A.frags[i];
    for (int i = threadIdx.x; i < kOut * kHid; i += kBlock) sw5[i] = A.params[oW5 + i];
    for (int i = threadIdx.x; i < 4 * kHid + kOut; i += kBlock)
        sbias[i] = A.params[i < 4 * kHid ? kBOff[i / kHid] + i % kHid : oB5 + (i - 4 * kHid)];
    __syncthreads();
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t nwaves = gridDim.x * (kBlock / 64);
    const DnFrameDev& F = A.F;
    const uint32_t npix = F.width * F.height;
    for (uint32_t base = (blockIdx.x * (kBlock / 64) + wave) * (16u * Q); base < A.nsamples;
         base += nwaves * (16u * Q)) {
        uint32_t px[Q], py[Q], pl[Q], pp[Q], row[Q];
        bool valid[Q], inb[Q];
        float dc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            uint32_t i = base + 16u * q + (uint32_t)j;
            uint32_t p = 0, l = 0;
            float dcv = 0.0f;
            bool in = i < A.nsamples, v = false;
            if (in) {
                uint32_t sm = A.samples[i];
                p = sm & 0xFFFFFFu;
                l = sm >> 24;
                // caller-supplied samples outside the frame contribute nothing
                if (p >= npix || l >= A.lights) { p = 0; l = 0; }
                else {
                    dcv = F.pix[p].x;
                    v = dcv >= 0.0f;
                }
            }
            pp[q] = p; pl[q] = l;
            px[q] = p % F.width; py[q] = p / F.width;
            row[q] = A.row_base + i;
            dc[q] = v ? dcv : 0.0f;
            valid[q] = v;
            inb[q] = in;
        }
        // ---- layer 1: 150 (160) -> 64
        f32x4 acc[Q][4];   // [q][t]
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[q][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 5; ++s) {
            bf16x8 b[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                uint4 fb = dn_features(F, valid[q], px[q], py[q], pl[q], dc[q], 32 * s + 8 * g);
                b[q] = as_bf16x8(fb);
                if (inb[q]) {
                    const uint32_t ws[4] = {fb.x, fb.y, fb.z, fb.w};
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        A.XT[(size_t)(32 * s + 8 * g + e) * A.ldn + row[q]] =
                            (uint16_t)(ws[e >> 1] >> (16 * (e & 1)));
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                bf16x8 a = as_bf16x8(sfrag[(kF1 + t * 5 + s) * 64 + lane]);
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q][t] = mfma(a, b[q], acc[q][t]);
            }
        }
        uint4 hb[Q][2];
        uint32_t mask[4][Q];   // [layer][q]
#pragma unroll
        for (int q = 0; q < Q; ++q)
            dn_hidden_epilogue(acc[q], sbias, g, hb[q], mask[0][q], A.HT, A.ldn, row[q], inb[q]);
        // ---- layers 2..4: 64 -> 64
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            dn_layer64<Q>(sfrag, k == 1 ? kF2 : (k == 2 ? kF3 : kF4), lane, hb, acc);
#pragma unroll
            for (int q = 0; q < Q; ++q)
                dn_hidden_epilogue(acc[q], sbias + k * kHid, g, hb[q], mask[k][q],
                                   A.HT + (size_t)k * kHid * A.ldn, A.ldn, row[q], inb[q]);
        }
        // ---- layer 5: 64 -> 2 in FP32 on layer 4's activations (as in dn_hidden_epilogue)
        float out[Q][2];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            f32x4 h[4];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float z = acc[q][t][i] + sbias[3 * kHid + 16 * t + 4 * g + i];
                    h[t][i] = z > 0.0f ? z : 0.0f;
                }
            dn_out_fp32(h, sw5, g, out[q][0], out[q][1]);
        }
        const float b50 = sbias[4 * kHid], b51 = sbias[4 * kHid + 1];
        // ---- loss and backward
        float loss = 0.0f;
        uint4 db[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            float d0 = 0.0f, d1 = 0.0f;
            if (g == 0 && valid[q]) {
                float bao, bvis;
                dn_box(F, px[q], py[q], pl[q], bao, bvis);
                float e0 = (out[q][0] + b50) + bao - A.target_ao[pp[q]];
                float e1 = (out[q][1] + b51) + bvis -
                           (F.num_lights ? A.target_vis[(size_t)pl[q] * npix + pp[q]] : 0.0f);
                loss += e0 * e0 + e1 * e1;
                d0 = A.grad_scale * e0;
                d1 = A.grad_scale * e1;
            }
            const uint32_t b01 = bf16_pack2(d0, d1);
            if (g == 0 && inb[q]) {
                A.YT[row[q]] = (uint16_t)b01;
                A.YT[(size_t)A.ldn + row[q]] = (uint16_t)(b01 >> 16);
            }
            db[q] = make_uint4(g == 0 ? b01 : 0u, 0u, 0u, 0u);
        }
        // dh4 = W5^T dy
#pragma unroll
        for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                acc[q][t] = mfma(as_bf16x8(sfrag[(kG5 + t) * 64 + lane]), as_bf16x8(db[q]),
                                 f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
        for (int k = 3; k >= 0; --k) {
            uint4 zb[Q][2];
#pragma unroll
            for (int q = 0; q < Q; ++q)
                dn_relu_back(acc[q], mask[k][q], g, zb[q], A.ZT + (size_t)k * kHid * A.ldn, A.ldn, row[q],
                             inb[q]);
            if (k == 0) break;
            // dh_{k} = W_{k+1}^T dz_{k+1}  (fragments G2..G4 hold W2^T..W4^T)
            dn_layer64<Q>(sfrag, k == 3 ? kG4 : (k == 2 ? kG3 : kG2), lane, zb, acc);
        }
        // per-wave loss sum (lanes g == 0 hold the outputs)
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
        if (lane == 0) atomicAdd(A.loss_sum, loss);
    }
}

struct InferArgs {
    DnFrameDev F;
    const uint4* frags;
    const float* params;
    uint32_t lights;    // max(L, 1)
    uint32_t ntiles;    // 16-pixel tiles of the frame
    float* ao_out;      // [npix]
    float* vis_out;     // [L][npix]
};

// Index of neighbour n (< 25) of pixel (x, y), clamped to the image.
__device__ inline uint32_t dn_nb_index(const DnFrameDev& F, uint32_t x, uint32_t y, int n) {
    int qx = (int)x + n % 5 - 2, qy = (int)y + n / 5 - 2;
    qx = qx < 0 ? 0 : (qx >= (int)F.width ? (int)F.width - 1 : qx);
    qy = qy < 0 ? 0 : (qy >= (int)F.height ? (int)F.height - 1 : qy);
    return (uint32_t)qy * F.width + (uint32_t)qx;
}

// The per-pixel state a tile's light groups share: lane (j, g) holds pixel j
// of the tile and neighbours 8g + e (e < 8) of its 5x5 window.
struct InferTile {
    uint32_t vq[8];   // neighbour indices (0 off the window and for invalid pixels)
    uint32_t bo[8];   // visibility table offset: 0 for a foreground neighbour, else
                      // 2048 bytes (the table's zero half)
    float cnt, bao;   // foreground pixels of the window, box base of AO
    uint32_t pp;      // pixel index (0 beyond the frame)
    bool valid, inb;
};

// Sum of a value over the four lane groups (rows of 16 lanes) of a pixel
// column, as ((g0 + g1) + (g2 + g3)), in every lane: v_permlane16_swap pairs
// rows 0/1 and 2/3, v_permlane32_swap the two halves (VALU exchanges, no LDS
// round trip).
__device__ inline float col_sum(float v) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    const float s = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(s), __float_as_uint(s), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}
__device__ inline uint32_t col_sum(uint32_t v) {
    const auto a = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    const uint32_t s = a[0] + a[1];
    const auto b = __builtin_amdgcn_permlane32_swap(s, s, false, false);
    return b[0] + b[1];
}

// Buffer resource over `bytes` bytes at p for raw loads (32-bit offsets in a
// VGPR, no 64-bit address arithmetic per load).
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}

// The prepass record (camera distance or -1, normal) of pixel q.
__device__ inline float4 load_pix(__amdgpu_buffer_rsrc_t r, uint32_t q) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, q * 16u, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}

// The shadow bytes of lights l0 .. l0+2 (clamped to the last light) at the
// tile's neighbour indices.
__device__ inline void dn_shadow_bytes(const DnFrameDev& F, uint32_t l0, const uint32_t (&vq)[8],
                                       uint32_t (&s8)[3][8]) {
    const uint32_t npix = F.width * F.height;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const uint32_t l = l0 + p < F.num_lights ? l0 + p : F.num_lights - 1;
        const __amdgpu_buffer_rsrc_t r = buf_rsrc(F.shadow + (size_t)l * npix, npix);
#pragma unroll
        for (int e = 0; e < 8; ++e) s8[p][e] = __builtin_amdgcn_raw_buffer_load_b8(r, vq[e], 0, 0);
    }
}

// Lights l0 .. l0+P-1 of one 16-pixel tile: the visibility K-step FV on top of
// the light-independent accumulators `ag`, then layers 2..5 on MFMA.  Every
// weight fragment and bias read from LDS feeds the products of all P lights;
// the caller loaded the group's shadow bytes `s8`.  AO outputs are added to
// `aosum` in light order; the visibilities are written.
//
// Layer 5 takes layer 4's activations in BF16, as every other layer takes its
// inputs, and W5 as the sum of two BF16 parts (F5), so only the activation
// rounding (relative 2^-9 per term) separates it from the FP32 layer 5 of
// training.
template <int P>
__device__ inline void dn_light_group(const InferArgs& A, const uint4* sfrag, const float* sbias,
                                      const uint2* svt, uint32_t l0, int lane, int g,
                                      const f32x4 (&ag)[4], const InferTile& T, const uint32_t (&s8)[3][8],
                                      float b50, float b51, float& aosum) {
    const DnFrameDev& F = A.F;
    const uint32_t npix = F.width * F.height;
    const uint32_t L = F.num_lights;
    // the weight fragments are read from LDS at every use: an opaque lane
    // offset keeps the compiler from holding all 28 of them in registers
    // across the groups (which would leave one wave per SIMD)
    uint32_t ln = (uint32_t)lane;
    asm volatile("" : "+v"(ln));
    f32x4 acc[P][4];
    uint4 hb[P][2];
    float bvis[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        float sv = 0.0f;
        if (L) {
            // (BF16 bits, FP32 bits) of 1 - s / shadow_n, or zeros (+0.0f, which
            // the sum keeps) for a neighbour off the foreground
            uint2 tv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e)
                tv[e] = *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(svt) +
                                                        ((s8[p][e] << 3) | T.bo[e]));
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                w[e >> 1] |= tv[e].x << (16 * (e & 1));
                sv += __uint_as_float(tv[e].y);
            }
        }
        bvis[p] = col_sum(sv) / T.cnt;
        hb[p][0] = make_uint4(w[0], w[1], w[2], w[3]);   // the visibility B operand
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const bf16x8 a = as_bf16x8(sfrag[(kIFV + t) * 64 + ln]);
#pragma unroll
        for (int p = 0; p < P; ++p) acc[p][t] = mfma(a, as_bf16x8(hb[p][0]), ag[t]);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) dn_relu_pack(acc[p], hb[p]);
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        const int fb = k == 1 ? kIF2 : (k == 2 ? kIF3 : kIF4);
        uint4 fr[8];   // fragment (t, s) at t * 2 + s, all 8 in flight together
#pragma unroll
        for (int i = 0; i < 8; ++i) fr[i] = sfrag[(fb + i) * 64 + ln];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const f32x4 bk = *reinterpret_cast<const f32x4*>(&sbias[k * kHid + 16 * t + 4 * g]);
#pragma unroll
            for (int p = 0; p < P; ++p) acc[p][t] = bk;
        }
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8 a = as_bf16x8(fr[t * 2 + s]);
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p][t] = mfma(a, as_bf16x8(hb[p][s]), acc[p][t]);
            }
#pragma unroll
        for (int p = 0; p < P; ++p) dn_relu_pack(acc[p], hb[p]);
    }
    // layer 5: lane group 0 receives rows 0..3 of its pixel column, the high
    // and low parts of the two outputs
    f32x4 y[P];
    {
        const bf16x8 a0 = as_bf16x8(sfrag[(kIF5 + 0) * 64 + ln]), a1 = as_bf16x8(sfrag[(kIF5 + 1) * 64 + ln]);
#pragma unroll
        for (int p = 0; p < P; ++p) y[p] = mfma(a0, as_bf16x8(hb[p][0]), f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
        for (int p = 0; p < P; ++p) y[p] = mfma(a1, as_bf16x8(hb[p][1]), y[p]);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        if (g == 0 && T.inb) {
            const float o0 = y[p][0] + y[p][2], o1 = y[p][1] + y[p][3];
            aosum += T.valid ? (o0 + b50) + T.bao : 0.0f;
            if (L)
                A.vis_out[(size_t)(l0 + p) * npix + T.pp] = T.valid ? clamp01f((o1 + b51) + bvis[p]) : 0.0f;
        }
    }
}

// Inference, one 16-pixel tile per wave at a time (lane column j: pixel j of
// the tile): the light-independent part of layer 1 once per pixel (K-steps FG,
// bias folded into the accumulators), then the lights in groups of 3 (the
// last group 1 or 2), each light with the visibility K-step FV on top of it
// and layers 2..5.  A pixel's AO is the mean of its lights' AO outputs, summed
// in light order; background pixels give 0.
__global__ __launch_bounds__(kInferBlock, 2) void k_dn_infer(InferArgs A) {
    __shared__ uint4 sfrag[kFragInfer * 64];
    __shared__ __align__(16) float sbias[4 * kHid + 4];
    __shared__ uint2 svt[512];   // (BF16 bits, FP32 bits) of 1 - s / shadow_n; zeros
    for (int i = threadIdx.x; i < 24 * 64; i += kInferBlock) sfrag[kIF2 * 64 + i] = A.frags[kF2 * 64 + i];
    for (int i = threadIdx.x; i < 24 * 64; i += kInferBlock) sfrag[kIFG * 64 + i] = A.frags[kFG * 64 + i];
    for (int i = threadIdx.x; i < 2 * 64; i += kInferBlock) sfrag[kIF5 * 64 + i] = A.frags[kF5 * 64 + i];
    for (int i = threadIdx.x; i < 4 * kHid + kOut; i += kInferBlock)
        sbias[i] = A.params[i < 4 * kHid ? kBOff[i / kHid] + i % kHid : oB5 + (i - 4 * kHid)];
    const DnFrameDev& F = A.F;
    for (int i = threadIdx.x; i < 256; i += kInferBlock) {
        const float v = 1.0f - (float)i / F.shadow_n;
        svt[i] = make_uint2(bf16_bits(v), __float_as_uint(v));
        svt[256 + i] = make_uint2(0u, 0u);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const uint32_t npix = F.width * F.height;
    const uint32_t L = F.num_lights;
    // The tiles go out in chunks of kInferRounds bands (a band: one tile per
    // wave of an XCD), chunk c to XCD c % 8 (blocks are dispatched round-robin
    // over the 8 XCDs): an XCD's waves sweep each of its chunks band after
    // band, so the window rows that neighbouring bands share are still in its
    // L2 when the next band reads them, and the chunks interleave the image
    // over the XCDs, which keeps their work even where whole regions are
    // background.  With a grid not a multiple of 8 every wave strides alone.
    const uint32_t nb = gridDim.x, wpb = kInferBlock / 64;
    const bool xcd = nb % 8 == 0;
    const uint32_t W = xcd ? (nb / 8) * wpb : nb * wpb;   // waves per band
    const uint32_t wx = (xcd ? blockIdx.x / 8 : blockIdx.x) * wpb + (threadIdx.x >> 6);
    const uint32_t c0 = xcd ? blockIdx.x % 8 : 0u, cstep = xcd ? 8u : 1u;
    const uint32_t C = kInferRounds * W;                   // tiles per chunk
    const float b50 = sbias[4 * kHid], b51 = sbias[4 * kHid + 1];
    const __amdgpu_buffer_rsrc_t rpix = buf_rsrc(F.pix, npix * 16u), rao = buf_rsrc(F.ao_frac, npix * 4u);
    for (uint32_t k = 0;; ++k) {
        const uint32_t base = (c0 + cstep * (k / kInferRounds)) * C;   // first tile of the chunk
        if (base >= A.ntiles) break;
        const uint32_t tile = base + (k % kInferRounds) * W + wx;
        if (tile >= A.ntiles) continue;   // the frame's last chunk ends early
        InferTile T;
        const uint32_t p = tile * 16u + (uint32_t)j;
        T.inb = p < npix;
        T.pp = T.inb ? p : 0u;
        const float d = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rpix, T.pp * 16u, 0, 0));
        T.valid = T.inb && d >= 0.0f;
        const float dc = T.valid ? d : 0.0f;
        const uint32_t px = T.pp % F.width, py = T.pp / F.width;
        if (__ballot(T.valid) == 0ull) {
            if (g == 0 && T.inb) {
                A.ao_out[T.pp] = 0.0f;
                for (uint32_t l = 0; l < L; ++l) A.vis_out[(size_t)l * npix + T.pp] = 0.0f;
            }
            continue;
        }
        // ---- layer 1, light-independent features: K-steps 0..3 (AO, normal)
        // of neighbours 8g + 2s + e / 4, then K-step 4: relative depth of
        // neighbours 8g + e, and the neighbour indices and foreground bits that
        // the visibility step reads per light.  Every neighbour index is
        // clamped to the image, so all loads are issued before the first is
        // used (invalid pixels and neighbours beyond the 25th select zeros).
        uint32_t nv[8];
        float4 vv[8];
        float av[8];
        bool in[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            nv[e] = dn_nb_index(F, px, py, 8 * g + e < 25 ? 8 * g + e : 0);
            in[e] = T.valid && 8 * g + e < 25;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            vv[e] = load_pix(rpix, nv[e]);
            av[e] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rao, nv[e] * 4u, 0, 0));
        }
        f32x4 ag[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) ag[t] = *reinterpret_cast<const f32x4*>(&sbias[16 * t + 4 * g]);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            uint32_t w[4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = 2 * s + h;
                const float4 v = in[e] ? vv[e] : make_float4(-1.0f, 0.0f, 0.0f, 0.0f);
                const float ao = v.x >= 0.0f ? av[e] : 0.0f;
                w[2 * h] = bf16_pack2(ao, v.y);     // background normals are zero
                w[2 * h + 1] = bf16_pack2(v.z, v.w);
            }
            const bf16x8 fb = as_bf16x8(make_uint4(w[0], w[1], w[2], w[3]));
#pragma unroll
            for (int t = 0; t < 4; ++t) ag[t] = mfma(as_bf16x8(sfrag[(kIFG + t * 5 + s) * 64 + lane]), fb, ag[t]);
        }
        {
            float f[8];
            float sa = 0.0f;
            uint32_t vfg = 0u;   // foreground bits of the neighbours 8g + e
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float dn = vv[e].x;
                const bool fg = in[e] && dn >= 0.0f;
                T.vq[e] = in[e] ? nv[e] : 0u;
                T.bo[e] = fg ? 0u : 2048u;
                vfg |= (fg ? 1u : 0u) << e;
                sa += fg ? av[e] : 0.0f;
                const float rel = dc > 0.0f ? (dn - dc) / dc * 16.0f : 0.0f;
                f[e] = !in[e] ? 0.0f : (!fg ? 1.0f : (rel < -1.0f ? -1.0f : (rel > 1.0f ? 1.0f : rel)));
            }
            const bf16x8 fb = as_bf16x8(make_uint4(bf16_pack2(f[0], f[1]), bf16_pack2(f[2], f[3]),
                                                   bf16_pack2(f[4], f[5]), bf16_pack2(f[6], f[7])));
            // sums over the four lane groups of a pixel column: ((g0 + g1) + (g2 + g3))
            sa = col_sum(sa);
            const uint32_t c = col_sum((uint32_t)__popc(vfg));
            T.cnt = (float)(c ? c : 1u);
            T.bao = sa / T.cnt;
#pragma unroll
            for (int t = 0; t < 4; ++t) ag[t] = mfma(as_bf16x8(sfrag[(kIFG + t * 5 + 4) * 64 + lane]), fb, ag[t]);
        }
        // ---- per light group: visibility K-step, layers 2..5
        float aosum = 0.0f;
        const uint32_t nl = A.lights;
        uint32_t sb[3][8];
        if (L) dn_shadow_bytes(F, 0, T.vq, sb);
        uint32_t l = 0;
        for (; l + 3 <= nl; l += 3) {
            if (L && l > 0) dn_shadow_bytes(F, l, T.vq, sb);
            dn_light_group<3>(A, sfrag, sbias, svt, l, lane, g, ag, T, sb, b50, b51, aosum);
        }
        if (L && l > 0 && l < nl) dn_shadow_bytes(F, l, T.vq, sb);
        if (nl - l == 2)
            dn_light_group<2>(A, sfrag, sbias, svt, l, lane, g, ag, T, sb, b50, b51, aosum);
        else if (nl - l == 1)
            dn_light_group<1>(A, sfrag, sbias, svt, l, lane, g, ag, T, sb, b50, b51, aosum);
        if (g == 0 && T.inb) A.ao_out[T.pp] = clamp01f(aosum / (float)nl);
    }
}

// dW[o][i] = sum_n dz[o][n] h[i][n] over one chunk of samples, for layer
// blockIdx.y (0..4); bias gradients = sum_n dz[o][n].  One partial per chunk.
struct WgArgs {
    const uint16_t* XT;
    const uint16_t* HT;
    const uint16_t* ZT;
    const uint16_t* YT;
    uint32_t ldn, rows;
    float* partial;   // [chunks][kParams]
};

__global__ __launch_bounds__(kBlock) void k_dn_wgrad(WgArgs A) {
    const int layer = blockIdx.y;
    const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
    const int wave = threadIdx.x >> 6;
    const uint16_t* Z = layer < 4 ? A.ZT + (size_t)layer * kHid * A.ldn : A.YT;
    const uint16_t* H = layer == 0 ? A.XT : A.HT + (size_t)(layer - 1) * kHid * A.ldn;
    const int Mt = layer < 4 ? 4 : 1, Kt = layer == 0 ? 10 : 4;
    const int nin = layer == 0 ? kIn : kHid, nout = layer < 4 ? kHid : kOut;
    const uint32_t n_begin = blockIdx.x * kChunkRows;
    const uint32_t n_end = n_begin + kChunkRows < A.rows ? n_begin + kChunkRows : A.rows;
    float* part = A.partial + (size_t)blockIdx.x * kParams;
    for (int tile = wave; tile < Mt * Kt; tile += kBlock / 64) {
        const int mt = tile / Kt, kt = tile % Kt;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const uint16_t* za = Z + (size_t)(16 * mt + j) * A.ldn + 8 * g;
        const uint16_t* hb = H + (size_t)(16 * kt + j) * A.ldn + 8 * g;
        for (uint32_t n0 = n_begin; n0 < n_end; n0 += 32) {
            uint4 a = *reinterpret_cast<const uint4*>(za + n0);
            uint4 b = *reinterpret_cast<const uint4*>(hb + n0);
            acc = mfma(as_bf16x8(a), as_bf16x8(b), acc);
        }
        const int in = 16 * kt + j;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = 16 * mt + 4 * g + i;
            if (o < nout && in < nin) part[kWOff[layer] + o * nin + in] = acc[i];
        }
    }
    // bias: four threads per output row, each a quarter of the chunk
    for (int r = threadIdx.x; r < nout * 4; r += kBlock) {
        const int o = r >> 2, part4 = r & 3;
        const uint16_t* zr = Z + (size_t)o * A.ldn;
        float s = 0.0f;
        for (uint32_t n = n_begin + part4 * 8; n < n_end; n += 32) {
            uint4 v = *reinterpret_cast<const uint4*>(zr + n);
            const uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 8; ++e) s += bf16_val((ws[e >> 1] >> (16 * (e & 1))) & 0xFFFFu);
        }
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (part4 == 0) part[kBOff[layer] + o] = s;
    }
}

__global__ void k_dn_reduce(const float* partial, uint32_t chunks, float* grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kParams) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) s += partial[(size_t)c * kParams + i];
    grad[i] = s;
}

__global__ void k_dn_adam(float* params, float* m, float* v, const float* grad, float lr_t, float b1,
                          float b2, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kParams) return;
    float gr = grad[i];
    float mi = b1 * m[i] + (1.0f - b1) * gr;
    float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// BF16 fragments of every weight matrix from the FP32 master copy.
__global__ void k_dn_pack(const float* P, uint16_t* frags) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (uint32_t)kFragAll * 64 * 8) return;
    const int f = id / 512, l = (id / 8) % 64, e = id % 8;
    const int j = l & 15, g = l >> 4;
    float val = 0.0f;
    if (f < kF2) {
        const int t = f / 5, s = f % 5, o = 16 * t + j, k = 32 * s + 8 * g + e;
        if (k < kIn) val = P[oW1 + o * kIn + k];
    } else if (f < kG2) {
        const int layer = 1 + (f - kF2) / 8, r = (f - kF2) % 8, t = r / 2, s = r % 2;
        val = P[kWOff[layer] + (16 * t + j) * kHid + khid(s, g, e)];
    } else if (f < kG5) {
        const int layer = 1 + (f - kG2) / 8, r = (f - kG2) % 8, t = r / 2, s = r % 2;
        val = P[kWOff[layer] + khid(s, g, e) * kHid + (16 * t + j)];
    } else if (f < kFG) {
        const int t = f - kG5, o = khid(0, g, e);
        if (o < kOut) val = P[oW5 + o * kHid + (16 * t + j)];
    } else if (f < kF5) {
        const int t = f < kFV ? (f - kFG) / 5 : f - kFV, s = f < kFV ? (f - kFG) % 5 : 5;
        const int k = infer_feature(s, g, e);
        if (k >= 0) val = P[oW1 + (16 * t + j) * kIn + k];
    } else if (j < 4) {
        const float w = P[oW5 + (j & 1) * kHid + khid(f - kF5, g, e)];
        const float hi = bf16_val(bf16_bits(w));
        val = j < 2 ? hi : w - hi;
    }
    frags[id] = (uint16_t)bf16_bits(val);
}

// Input features of explicit samples, [count][150] BF16 bits (inspection).
__global__ __launch_bounds__(kBlock) void k_dn_features(DnFrameDev F, const uint32_t* samples,
                                                        uint32_t count, uint16_t* out) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t i = id / 19, grp = id % 19;   // 19 groups of 8 cover 150 features
    if (i >= count) return;
    uint32_t sm = samples[i], p = sm & 0xFFFFFFu, l = sm >> 24;
    const uint32_t lights = F.num_lights ? F.num_lights : 1u;
    bool valid = p < F.width * F.height && l < lights;
    if (!valid) { p = 0; l = 0; }
    const float dcv = F.pix[p].x;
    valid = valid && dcv >= 0.0f;
    const float dc = valid ? dcv : 0.0f;
    uint4 fb = dn_features(F, valid, p % F.width, p / F.width, l, dc, 8 * (int)grp);
    const uint32_t ws[4] = {fb.x, fb.y, fb.z, fb.w};
    for (int e = 0; e < 8; ++e) {
        int f = 8 * (int)grp + e;
        if (f < kIn) out[(size_t)i * kIn + f] = (uint16_t)(ws[e >> 1] >> (16 * (e & 1)));
    }
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace
}  // namespace luz

using namespace luz;

struct luzrt_denoiser {
    uint32_t capacity = 0;     // training samples per gradient
    uint32_t rows = 0;         // samples accumulated since the last gradient
    uint32_t step = 0;         // Adam steps taken
    float* params = nullptr;   // FP32 master
    float* grad = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* partial = nullptr;
    float* loss = nullptr;     // device loss sum
    uint16_t* frags = nullptr;
    uint16_t* XT = nullptr;
    uint16_t* HT = nullptr;
    uint16_t* ZT = nullptr;
    uint16_t* YT = nullptr;
    DevBuf pix, ao_frac;    // prepass planes of the current frame
};

namespace {

void dn_free(luzrt_denoiser* d) {
    for (void* p : {(void*)d->params, (void*)d->grad, (void*)d->adam_m, (void*)d->adam_v,
                    (void*)d->partial, (void*)d->loss, (void*)d->frags, (void*)d->XT, (void*)d->HT,
                    (void*)d->ZT, (void*)d->YT, d->pix.ptr, d->ao_frac.ptr})
        if (p) (void)hipFree(p);
    delete d;
}

luzrt_status frame_dev(luzrt_ctx* ctx, const luzrt_dn_frame* f, DnFrameDev& F) {
    if (!f || !f->position || !f->normal || !f->ao_counts || (f->num_lights && !f->shadow_counts))
        return fail(ctx, LUZRT_E_INVALID_ARG, "denoiser frame: null plane");
    if (f->width == 0 || f->height == 0 || (uint64_t)f->width * f->height > 0xFFFFFFull)
        return fail(ctx, LUZRT_E_LIMIT, "denoiser frame: %ux%u outside 1..2^24 pixels", f->width, f->height);
    if (f->num_lights > LUZRT_MAX_LIGHTS || f->ao_samples == 0 || (f->num_lights && f->shadow_samples == 0))
        return fail(ctx, LUZRT_E_INVALID_ARG, "denoiser frame: bad light or sample counts");
    std::memset(&F, 0, sizeof(F));
    F.width = f->width;
    F.height = f->height;
    F.num_lights = f->num_lights;
    F.position = f->position;
    F.normal = f->normal;
    F.shadow = f->shadow_counts;
    F.ao = f->ao_counts;
    for (int k = 0; k < 3; ++k) F.cam[k] = f->cam_pos[k];
    F.shadow_n = (float)(f->shadow_samples ? f->shadow_samples : 1u);
    F.ao_n = (float)f->ao_samples;
    return LUZRT_OK;
}

// Runs the per-pixel prepass of a frame into the given planes.
luzrt_status prep_frame(luzrt_ctx* ctx, DnFrameDev& F, DevBuf& pix, DevBuf& ao_frac) {
    const uint32_t npix = F.width * F.height;
    luzrt_status s;
    if ((s = ensure(ctx, pix, (size_t)npix * 16)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ao_frac, (size_t)npix * 4)) != LUZRT_OK) return s;
    k_dn_prep<<<cdiv(npix, 256), 256, 0, ctx->stream>>>(F, (float4*)pix.ptr, (float*)ao_frac.ptr);
    LUZ_HIP(ctx, hipGetLastError());
    F.pix = (const float4*)pix.ptr;
    F.ao_frac = (const float*)ao_frac.ptr;
    return LUZRT_OK;
}

luzrt_status repack(luzrt_ctx* ctx, luzrt_denoiser* d) {
    const uint32_t n = kFragAll * 64 * 8;
    k_dn_pack<<<cdiv(n, 256), 256, 0, ctx->stream>>>(d->params, d->frags);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

}  // namespace

extern "C" {

luzrt_status luzrt_denoiser_create(luzrt_ctx* ctx, uint32_t batch_capacity, uint64_t seed,
                                   luzrt_denoiser** out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!out) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoiser_create: null argument");
    *out = nullptr;
    if (batch_capacity % 64 != 0 || batch_capacity > (1u << 22))
        return fail(ctx, LUZRT_E_LIMIT, "batch capacity %u: a multiple of 64 up to 2^22", batch_capacity);
    LUZ_HIP(ctx, hipSetDevice(ctx->device));
    auto* d = new luzrt_denoiser();
    d->capacity = batch_capacity;
    const uint32_t cap = batch_capacity ? batch_capacity : 64;
    const uint32_t chunks = cdiv(cap, kChunkRows);
    const size_t pb = (size_t)kParams * 4;
    struct A { void** p; size_t bytes; } allocs[] = {
        {(void**)&d->params, pb}, {(void**)&d->grad, pb}, {(void**)&d->adam_m, pb},
        {(void**)&d->adam_v, pb}, {(void**)&d->partial, pb * chunks}, {(void**)&d->loss, 256},
        {(void**)&d->frags, (size_t)kFragAll * 1024}, {(void**)&d->XT, (size_t)160 * cap * 2},
        {(void**)&d->HT, (size_t)4 * kHid * cap * 2}, {(void**)&d->ZT, (size_t)4 * kHid * cap * 2},
        {(void**)&d->YT, (size_t)16 * cap * 2}};
    for (auto& a : allocs) {
        hipError_t e = hipMalloc(a.p, a.bytes);
        if (e != hipSuccess) { dn_free(d); return hip_fail(ctx, e, "hipMalloc(denoiser)"); }
    }
    // rows 2..15 of dy are never written: zero them once
    LUZ_HIP(ctx, hipMemset(d->YT, 0, (size_t)16 * cap * 2));
    LUZ_HIP(ctx, hipMemset(d->adam_m, 0, pb));
    LUZ_HIP(ctx, hipMemset(d->adam_v, 0, pb));
    LUZ_HIP(ctx, hipMemset(d->grad, 0, pb));
    LUZ_HIP(ctx, hipMemset(d->loss, 0, 256));
    // He-uniform initial weights (SplitMix64), zero biases
    std::vector<float> h(kParams, 0.0f);
    uint64_t st = seed;
    auto next = [&st]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    const int fan_in[5] = {kIn, kHid, kHid, kHid, kHid}, rows_out[5] = {kHid, kHid, kHid, kHid, kOut};
    for (int L = 0; L < 5; ++L) {
        double lim = std::sqrt(6.0 / fan_in[L]);
        for (int i = 0; i < rows_out[L] * fan_in[L]; ++i) h[kWOff[L] + i] = (float)((2.0 * next() - 1.0) * lim);
    }
    LUZ_HIP(ctx, hipMemcpy(d->params, h.data(), pb, hipMemcpyHostToDevice));
    luzrt_status s = repack(ctx, d);
    if (s != LUZRT_OK) { dn_free(d); return s; }
    *out = d;
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_destroy(luzrt_ctx* ctx, luzrt_denoiser* d) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d) return LUZRT_OK;
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dn_free(d);
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_get_params(luzrt_ctx* ctx, luzrt_denoiser* d, float* host, uint32_t count) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !host || count != (uint32_t)kParams)
        return fail(ctx, LUZRT_E_INVALID_ARG, "denoiser parameters: need %d floats", kParams);
    LUZ_HIP(ctx, hipMemcpyAsync(host, d->params, (size_t)kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_set_params(luzrt_ctx* ctx, luzrt_denoiser* d, const float* host, uint32_t count) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !host || count != (uint32_t)kParams)
        return fail(ctx, LUZRT_E_INVALID_ARG, "denoiser parameters: need %d floats", kParams);
    for (uint32_t i = 0; i < count; ++i)
        if (!std::isfinite(host[i])) return fail(ctx, LUZRT_E_INVALID_ARG, "parameter %u is not finite", i);
    LUZ_HIP(ctx, hipMemcpyAsync(d->params, host, (size_t)kParams * 4, hipMemcpyHostToDevice, ctx->stream));
    LUZ_HIP(ctx, hipMemsetAsync(d->adam_m, 0, (size_t)kParams * 4, ctx->stream));
    LUZ_HIP(ctx, hipMemsetAsync(d->adam_v, 0, (size_t)kParams * 4, ctx->stream));
    d->step = 0;
    luzrt_status s = repack(ctx, d);
    if (s != LUZRT_OK) return s;
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

luzrt_status luzrt_denoise(luzrt_ctx* ctx, luzrt_denoiser* d, const luzrt_dn_frame* f, float* ao_out,
                           float* vis_out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !ao_out || (f && f->num_lights && !vis_out))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoise: null argument");
    InferArgs A{};
    luzrt_status s = frame_dev(ctx, f, A.F);
    if (s != LUZRT_OK) return s;
    const uint32_t npix = f->width * f->height;
    stage_begin(ctx, kStageDenoise);
    if ((s = prep_frame(ctx, A.F, d->pix, d->ao_frac)) != LUZRT_OK) return s;
    A.frags = reinterpret_cast<const uint4*>(d->frags);
    A.params = d->params;
    A.lights = f->num_lights ? f->num_lights : 1u;
    A.ntiles = cdiv(npix, 16);
    A.ao_out = ao_out;
    A.vis_out = vis_out;
    // as many groups as are resident at once (48 KB of LDS and the registers
    // of 8 waves each), so that no group waits for another to finish
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dn_infer, kInferBlock, 0) !=
            hipSuccess || per_cu < 1)
        per_cu = 1;
    uint32_t groups = cdiv(A.ntiles, kInferBlock / 64);
    const uint32_t maxg = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 1) * (uint32_t)per_cu;
    if (groups > maxg) groups = maxg;
    k_dn_infer<<<groups, kInferBlock, 0, ctx->stream>>>(A);
    stage_end(ctx, kStageDenoise);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_features(luzrt_ctx* ctx, const luzrt_dn_frame* f, const uint32_t* samples,
                                     uint32_t count, uint16_t* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!samples || !out) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoiser_features: null argument");
    DnFrameDev F;
    luzrt_status s = frame_dev(ctx, f, F);
    if (s != LUZRT_OK) return s;
    if (count == 0) return LUZRT_OK;
    DevBuf pix, ao_frac;
    s = prep_frame(ctx, F, pix, ao_frac);
    if (s == LUZRT_OK) {
        k_dn_features<<<cdiv((uint64_t)count * 19, kBlock), kBlock, 0, ctx->stream>>>(F, samples, count, out);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) s = hip_fail(ctx, e, "k_dn_features");
    }
    // the planes are released once the kernel has read them
    (void)hipStreamSynchronize(ctx->stream);
    if (pix.ptr) (void)hipFree(pix.ptr);
    if (ao_frac.ptr) (void)hipFree(ao_frac.ptr);
    return s;
}

luzrt_status luzrt_denoiser_accumulate(luzrt_ctx* ctx, luzrt_denoiser* d, const luzrt_dn_frame* f,
                                       const float* target_ao, const float* target_vis,
                                       const uint32_t* samples, uint32_t count, float grad_scale) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !target_ao || !samples || (f && f->num_lights && !target_vis))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoiser_accumulate: null argument");
    if (count % 64 != 0 || d->rows + count > d->capacity)
        return fail(ctx, LUZRT_E_LIMIT, "%u samples (a multiple of 64) over the batch capacity %u (%u used)",
                    count, d->capacity, d->rows);
    DnArgs A{};
    luzrt_status s = frame_dev(ctx, f, A.F);
    if (s != LUZRT_OK) return s;
    if (count == 0) return LUZRT_OK;
    if ((s = prep_frame(ctx, A.F, d->pix, d->ao_frac)) != LUZRT_OK) return s;
    A.frags = reinterpret_cast<const uint4*>(d->frags);
    A.params = d->params;
    A.nsamples = count;
    A.lights = f->num_lights ? f->num_lights : 1u;
    A.samples = samples;
    A.target_ao = target_ao;
    A.target_vis = target_vis;
    A.row_base = d->rows;
    A.ldn = d->capacity;
    A.XT = d->XT;
    A.HT = d->HT;
    A.ZT = d->ZT;
    A.YT = d->YT;
    A.grad_scale = grad_scale;
    A.loss_sum = d->loss;
    uint32_t groups = cdiv(count, kBlock);
    const uint32_t maxg = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 1) * 4;
    if (groups > maxg) groups = maxg;
    k_dn_forward<4><<<groups, kBlock, 0, ctx->stream>>>(A);
    LUZ_HIP(ctx, hipGetLastError());
    d->rows += count;
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_gradient(luzrt_ctx* ctx, luzrt_denoiser* d, float* host_grad, float* loss_sum) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoiser_gradient: null denoiser");
    hipStream_t st = ctx->stream;
    if (d->rows > 0) {
        WgArgs W{d->XT, d->HT, d->ZT, d->YT, d->capacity, d->rows, d->partial};
        const uint32_t chunks = cdiv(d->rows, kChunkRows);
        k_dn_wgrad<<<dim3(chunks, 5), kBlock, 0, st>>>(W);
        k_dn_reduce<<<cdiv(kParams, 256), 256, 0, st>>>(d->partial, chunks, d->grad);
        LUZ_HIP(ctx, hipGetLastError());
    } else {
        LUZ_HIP(ctx, hipMemsetAsync(d->grad, 0, (size_t)kParams * 4, st));
    }
    d->rows = 0;
    if (host_grad) LUZ_HIP(ctx, hipMemcpyAsync(host_grad, d->grad, (size_t)kParams * 4, hipMemcpyDeviceToHost, st));
    if (loss_sum) LUZ_HIP(ctx, hipMemcpyAsync(loss_sum, d->loss, 4, hipMemcpyDeviceToHost, st));
    LUZ_HIP(ctx, hipMemsetAsync(d->loss, 0, 4, st));
    if (host_grad || loss_sum) LUZ_HIP(ctx, hipStreamSynchronize(st));
    return LUZRT_OK;
}

luzrt_status luzrt_denoiser_adam(luzrt_ctx* ctx, luzrt_denoiser* d, float lr, float beta1, float beta2,
                                 float eps) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!d || !(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_denoiser_adam: bad arguments");
    d->step++;
    // bias-corrected step size
    double c1 = 1.0 - std::pow((double)beta1, (double)d->step);
    double c2 = 1.0 - std::pow((double)beta2, (double)d->step);
    float lr_t = (float)(lr * std::sqrt(c2) / c1);
    k_dn_adam<<<cdiv(kParams, 256), 256, 0, ctx->stream>>>(d->params, d->adam_m, d->adam_v, d->grad, lr_t,
                                                          beta1, beta2, eps);
    LUZ_HIP(ctx, hipGetLastError());
    return repack(ctx, d);
}

}  // extern "C"
