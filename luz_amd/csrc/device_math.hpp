// device_math.hpp — arithmetic shared by the gfx950 kernels of the lighting path.
//
// Every FP32 expression here is evaluated exactly as written: the library is
// compiled with -ffp-contract=off (no fused multiply-add), IEEE division and
// square root (HIP's default correctly rounded forms) and denormals kept.
// Square roots use sqrtf: on gfx950 it lowers to v_sqrt_f32 plus a
// correction step (correctly rounded), whereas __fsqrt_rn lowers to the bare
// v_sqrt_f32 approximation.
// The shadow mask and AO counts are therefore reproducible bit for bit by a
// CPU implementation of the same specification.
//
// Specification (the same one the CPU oracle restates):
//  S1 samples: PCG4D hash (Jarzynski & Olano 2020) of (x, y, frame, slot),
//     u = (h.x >> 8) * 2^-24, v = (h.y >> 8) * 2^-24; shadow sample s of light
//     l uses slot l*256+s, AO sample s uses slot 16384+s.  Replaces the
//     blue-noise lookup of light.frag:71-75 (the texture asset is missing).
//  S2 sincos(2*pi*w): quadrant from floor(4w), Taylor polynomials of
//     sin/cos(pi/2 f) in Horner form.
//  S3 watertight triangle test (Woop, Benthin, Wald 2013), FP64 edge
//     fallback, no culling (VulkanWrapper.cpp:1119), closed [tmin, tmax].
//  S4 slab test: near/far plane by sign of 1/d, NaN-ignoring max/min,
//     interval [tmin - m, tmax + m] with m = max(1,|tmax|)*2^-10, far bound
//     widened by |tfar|*2^-20 (conservative against S3).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace luz {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFEu;

// 64-byte binary BVH node: slab bounds of both children, grouped per axis so
// one 16-byte load gives (lo0, lo1, hi0, hi1) of an axis.  Children with
// bit 31 set are leaves (BLAS: triangle slot; TLAS: instance slot);
// kEmpty marks an absent child (its box is inverted and never hit).
struct alignas(16) Node {
    float lox[2], hix[2];
    float loy[2], hiy[2];
    float loz[2], hiz[2];
    uint32_t child[2];
    uint32_t parent;
    uint32_t pad;
};
static_assert(sizeof(Node) == 64, "node is 64 bytes");

// TLAS leaf record, 64 bytes: world->object transform and the BLAS arrays.
struct alignas(16) InstRec {
    float w2o[12];
    const Node* nodes;
    const float* tris;
};
static_assert(sizeof(InstRec) == 64, "instance record is 64 bytes");

// 8-wide node of the precision form, 96 bytes: a per-axis power-of-two
// frame (origin, exponent) and 8-bit child bounds rounded outward, so the
// decoded child box origin + q * 2^e is exact in FP32 and contains the child.
// meta[k]: bit 7 set = internal child child_base + (meta & 0x7f), clear =
// triangle tri_base + meta.  Slots >= nchild are unused.
struct alignas(16) Node8 {
    float ox, oy, oz;
    int8_t ex, ey, ez;
    uint8_t nchild;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
    uint32_t pad[4];
};
static_assert(sizeof(Node8) == 96, "wide node is 96 bytes");

// Per-instance data used only by the primary-ray producer.
struct InstAux {
    float nmat[9];            // inverse-transpose of the 3x3 part, row-major
    uint32_t custom_index;
    uint32_t input_index;
    uint32_t pad;
    const float* nrms;        // per-slot vertex normals (9 floats)
    const uint32_t* prims;    // slot -> primitive id
};

struct V3 { float x, y, z; };

__host__ __device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__host__ __device__ inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 vcross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float vlen(V3 a) { return sqrtf(vdot(a, a)); }
// normalize() as the oracle restates GLSL's: the correctly rounded
// reciprocal of the correctly rounded length, then three products (one
// division instead of three).
__device__ inline V3 vnorm(V3 a) {
    const float r = __fdiv_rn(1.0f, vlen(a));
    return v3(a.x * r, a.y * r, a.z * r);
}
__device__ inline float fmin_(float a, float b) { return a < b ? a : b; }
__device__ inline float fmax_(float a, float b) { return a > b ? a : b; }

// ------------------------------------------------------------------ S1/S2
__device__ inline void pcg4d(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    a = a * 1664525u + 1013904223u; b = b * 1664525u + 1013904223u;
    c = c * 1664525u + 1013904223u; d = d * 1664525u + 1013904223u;
    a += b * d; b += c * a; c += a * b; d += b * c;
    a ^= a >> 16; b ^= b >> 16; c ^= c >> 16; d ^= d >> 16;
    a += b * d; b += c * a; c += a * b; d += b * c;
}

__device__ inline void sample2(uint32_t x, uint32_t y, uint32_t frame, uint32_t slot, float& u,
                               float& v) {
    uint32_t a = x, b = y, c = frame, d = slot;
    pcg4d(a, b, c, d);
    u = (float)(a >> 8) * 0x1p-24f;
    v = (float)(b >> 8) * 0x1p-24f;
}

__device__ inline void sincos_turns(float w, float& s, float& c) {
    float w4 = w * 4.0f;
    float q = floorf(w4);
    float f = w4 - q;
    float f2 = f * f;
    float ps = 5.692172921967926e-08f;
    ps = ps * f2; ps = ps + -3.598843235212085e-06f;
    ps = ps * f2; ps = ps + 0.00016044118478735982f;
    ps = ps * f2; ps = ps + -0.004681754135318687f;
    ps = ps * f2; ps = ps + 0.07969262624616703f;
    ps = ps * f2; ps = ps + -0.6459640975062462f;
    ps = ps * f2; ps = ps + 1.5707963267948966f;
    float qs = ps * f;
    float pc = -6.386603083791069e-09f;
    pc = pc * f2; pc = pc + 4.710874778818171e-07f;
    pc = pc * f2; pc = pc + -2.5202042373060604e-05f;
    pc = pc * f2; pc = pc + 0.0009192602748394265f;
    pc = pc * f2; pc = pc + -0.020863480763352957f;
    pc = pc * f2; pc = pc + 0.2536695079010480f;
    pc = pc * f2; pc = pc + -1.2337005501361697f;
    pc = pc * f2; pc = pc + 1.0f;
    float qc = pc;
    int qi = (int)q & 3;
    s = qi == 0 ? qs : (qi == 1 ? qc : (qi == 2 ? -qs : -qc));
    c = qi == 0 ? qc : (qi == 1 ? -qs : (qi == 2 ? -qc : qs));
}

// light.frag:65 uses 6.283*v radians; in turns that is v * 6.283/(2*pi).
constexpr float kHemiTurns = 0.9999705074463785f;

// ------------------------------------------------------------------ S3
struct TriRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ inline float sel3(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

__device__ inline TriRay tri_setup(V3 d) {
    TriRay r;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = sel3(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    // shear constants from one correctly rounded reciprocal (any values
    // shared by all triangles of the ray keep the test watertight)
    r.Sz = __fdiv_rn(1.0f, dz);
    r.Sx = sel3(d, kx) * r.Sz;
    r.Sy = sel3(d, ky) * r.Sz;
    return r;
}

// a*b - c*d to within 2 ulps (Kahan's algorithm with fused multiply-adds;
// Jeannerod, Louvet and Muller, Math. Comp. 82 (2013)): e is the exact
// rounding error of c*d, so the result is 0 only when a*b == c*d exactly and
// otherwise has the sign of the exact difference (products away from the
// FP32 underflow range).  The watertight test's edge functions fall back to
// it where the plain FP32 differences give 0, in place of the published FP64
// re-evaluation (Woop, Benthin and Wald, JCGT 2(1), 2013, section 3.3).
__device__ inline float diff_of_products(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = __builtin_fmaf(-c, d, w);
    const float f = __builtin_fmaf(a, b, -w);
    return f + e;
}

// Returns true on a hit with T/det in [tmin, tmax]; signed T, det, U, V, W out.
__device__ inline bool tri_test(const TriRay& r, V3 o, float tmin, float tmax, V3 v0, V3 v1,
                                V3 v2, float& T_out, float& det_out, float& U_out, float& V_out,
                                float& W_out) {
    V3 A = vsub(v0, o), B = vsub(v1, o), C = vsub(v2, o);
    float Akz = sel3(A, r.kz), Bkz = sel3(B, r.kz), Ckz = sel3(C, r.kz);
    float Ax = sel3(A, r.kx) - r.Sx * Akz;
    float Ay = sel3(A, r.ky) - r.Sy * Akz;
    float Bx = sel3(B, r.kx) - r.Sx * Bkz;
    float By = sel3(B, r.ky) - r.Sy * Bkz;
    float Cx = sel3(C, r.kx) - r.Sx * Ckz;
    float Cy = sel3(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = diff_of_products(Cx, By, Cy, Bx);
        V = diff_of_products(Ax, Cy, Ay, Cx);
        W = diff_of_products(Bx, Ay, By, Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float adet = det < 0.0f ? -det : det;
    float aT = det < 0.0f ? -T : T;
    if (aT < tmin * adet || aT > tmax * adet) return false;
    T_out = T; det_out = det; U_out = U; V_out = V; W_out = W;
    return true;
}

// ------------------------------------------------------------------ S4
__device__ inline float t_margin(float tmax) {
    float a = fabsf(tmax);
    return (a > 1.0f ? a : 1.0f) * 0x1p-10f;
}

struct BoxRay {
    V3 o, inv;
    float tlo, thi;
};

// Slab reciprocal by the hardware reciprocal (within 1 ulp, signed zeros to
// signed infinities).  The box test only has to be conservative and its far
// bound is widened by |tfar|·2^-20, eight ulps, so the reciprocal need not be
// correctly rounded.  v_rcp_f32 flushes denormal inputs: the component is
// scaled by 2^24 first (exact below 2^104) and the result scaled back.
__device__ inline float slab_rcp(float d) {
    return __builtin_amdgcn_rcpf(d * 0x1p24f) * 0x1p24f;
}

__device__ inline BoxRay box_setup(V3 o, V3 d, float tmin, float tmax) {
    BoxRay b;
    b.o = o;
    b.inv = v3(slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z));
    float m = t_margin(tmax);
    b.tlo = tmin - m;
    b.thi = tmax + m;
    return b;
}

// Slab test of child c of a node given its per-axis (lo0, lo1, hi0, hi1).
// The slab test on the box's near and far planes per axis (near: the lower
// plane where the ray's reciprocal direction is >= 0, else the upper one).
__device__ inline bool box_hit_nf(const BoxRay& r, float nx, float fx, float ny, float fy, float nz, float fz,
                                  float& tnear) {
    float tnx = (nx - r.o.x) * r.inv.x, tfx = (fx - r.o.x) * r.inv.x;
    float tny = (ny - r.o.y) * r.inv.y, tfy = (fy - r.o.y) * r.inv.y;
    float tnz = (nz - r.o.z) * r.inv.z, tfz = (fz - r.o.z) * r.inv.z;
    // max/min ignore a NaN slab distance (a zero direction component on a slab
    // plane) as the comparisons of the CPU restatement do
    const float tn = fmaxf(fmaxf(fmaxf(r.tlo, tnx), tny), tnz);
    float tf = fminf(fminf(fminf(r.thi, tfx), tfy), tfz);
    tf = tf + fabsf(tf) * 0x1p-20f;
    tnear = tn;
    return tn <= tf;
}

// The same test from the slab distances themselves (near and far per axis).
__device__ inline bool box_hit_t(const BoxRay& r, float tnx, float tfx, float tny, float tfy, float tnz, float tfz) {
    const float tn = fmaxf(fmaxf(fmaxf(r.tlo, tnx), tny), tnz);
    float tf = fminf(fminf(fminf(r.thi, tfx), tfy), tfz);
    tf = tf + fabsf(tf) * 0x1p-20f;
    return tn <= tf;
}

__device__ inline bool box_hit(const BoxRay& r, float lox, float hix, float loy, float hiy,
                               float loz, float hiz, float& tnear) {
    bool sx = r.inv.x < 0.0f, sy = r.inv.y < 0.0f, sz = r.inv.z < 0.0f;
    return box_hit_nf(r, sx ? hix : lox, sx ? lox : hix, sy ? hiy : loy, sy ? loy : hiy, sz ? hiz : loz,
                      sz ? loz : hiz, tnear);
}

__device__ inline V3 xform_point(const float* w, V3 p) {
    return v3(w[0] * p.x + w[1] * p.y + w[2] * p.z + w[3],
              w[4] * p.x + w[5] * p.y + w[6] * p.z + w[7],
              w[8] * p.x + w[9] * p.y + w[10] * p.z + w[11]);
}
__device__ inline V3 xform_vec(const float* w, V3 v) {
    return v3(w[0] * v.x + w[1] * v.y + w[2] * v.z,
              w[4] * v.x + w[5] * v.y + w[6] * v.z,
              w[8] * v.x + w[9] * v.y + w[10] * v.z);
}

}  // namespace luz
