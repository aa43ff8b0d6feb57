/*
 * luz_oracle.cpp — CPU oracle for the Luz ray-traced lighting path.
 *
 * TEST INFRASTRUCTURE ONLY (see luz_oracle.h).  Built with g++ -O3
 * -ffp-contract=off and no -ffast-math, so every FP32 operation below is a
 * single IEEE-754 round-to-nearest operation, evaluated in the order written.
 * The HIP kernels implement the same arithmetic specification independently;
 * tests compare the two bit for bit.
 *
 * Arithmetic specification (shared with the HIP path, restated there):
 *   S1  PCG4D hash (Jarzynski & Olano 2020, JCGT 9(3)) of (x, y, frame, slot);
 *       u = (h.x >> 8) * 2^-24, v = (h.y >> 8) * 2^-24.  Shadow sample s of
 *       light l uses slot l*256+s, AO sample s uses slot 16384+s.  This
 *       replaces light.frag:71-75 (blue-noise texture, asset missing).
 *   S2  sincos of 2*pi*w by quadrant reduction (w*4 split into integer and
 *       exact fraction) and Taylor polynomials of sin/cos(pi/2 f) evaluated
 *       by Horner with separate multiply and add.
 *   S3  Watertight ray/triangle test (Woop, Benthin, Wald 2013, JCGT 2(1)),
 *       double-precision edge-function fallback, no culling, closed interval
 *       [tmin, tmax].
 *   S4  Slab test with near/far plane chosen by the sign of 1/d, NaN-ignoring
 *       max/min, interval [tmin - m, tmax + m], m = max(1,|tmax|) * 2^-10, and
 *       far bound widened by |tfar| * 2^-20.
 *   S5  LBVH (Karras 2012): triangle AABBs widened by max|coord| * 2^-16,
 *       FP64 centroid frame, 10-bit-per-axis Morton codes, stable sort by
 *       (code, primitive index), Karras split with index tie-break, exact
 *       min/max refit; binary 64-byte nodes; one primitive per leaf.
 *   S6  Instances: world->object = FP64 inverse of the FP32 object->world,
 *       rounded to FP32; instance world boxes from the 8 transformed corners
 *       of the BLAS root box, widened by (max |coord| over all instances) *
 *       2^-14.
 *   S7  TLAS hierarchy: the S5 Morton slot order, then locally-ordered
 *       clustering (PLOC, Meister & Bittner 2018) with radius 16 and the
 *       FP32 union-box half-area as merge cost (build_ploc below).
 *
 * Parity pins (tests/test_oracle_*.py): Morton and PCG are checked against
 * independent brute-force/bit-level restatements and statistical properties;
 * the triangle test against closed-form barycentric cases and watertightness
 * on shared edges and vertices; the BVH against brute force over all
 * triangles; the TLAS clustering against exhaustive greedy agglomerative
 * clustering (which PLOC equals when the radius spans every cluster) and a
 * closed-form merge order; sincos against numpy; the instance inverse against the identity;
 * the GGX terms against closed forms (Walter et al. 2007, Karis 2013).
 * The Vulkan driver's own builder, traversal order and interval closure
 * cannot be run here (no Vulkan, closed driver): parity unpinned for those.
 */
#include "luz_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kLeaf = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kSentinel = 0xFFFFFFFEu;
constexpr float kPi = 3.14159265359f;  // LuzCommon.h:10

struct f3 { float x, y, z; };
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
// GLSL's normalize restated: one correctly rounded reciprocal of the length,
// then three products.
inline f3 normalize(f3 a) { float r = 1.0f / length(a); return mk(a.x * r, a.y * r, a.z * r); }
inline float fmin_(float a, float b) { return a < b ? a : b; }
inline float fmax_(float a, float b) { return a > b ? a : b; }
inline float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline f3 load3(const float* p) { return mk(p[0], p[1], p[2]); }

struct d3 { double x, y, z; };
inline d3 dk(double x, double y, double z) { return d3{x, y, z}; }
inline d3 dadd(d3 a, d3 b) { return dk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline d3 dsub(d3 a, d3 b) { return dk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline d3 dmul(d3 a, double s) { return dk(a.x * s, a.y * s, a.z * s); }
inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 dcross(d3 a, d3 b) {
    return dk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double dlength(d3 a) { return std::sqrt(ddot(a, a)); }
inline d3 dnormalize(d3 a) { double l = dlength(a); return dk(a.x / l, a.y / l, a.z / l); }
inline d3 promote(f3 a) { return dk(a.x, a.y, a.z); }
inline double dcomp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------- S1 / S2
inline void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}

inline void sample2(uint32_t x, uint32_t y, uint32_t frame, uint32_t slot, float& u, float& v) {
    uint32_t h[4] = {x, y, frame, slot};
    pcg4d(h);
    u = (float)(h[0] >> 8) * 0x1p-24f;
    v = (float)(h[1] >> 8) * 0x1p-24f;
}

// sin(pi/2 f), cos(pi/2 f) for f in [0,1): Taylor coefficients (pi/2)^k/k!.
inline void sincos_quarter(float f, float& s, float& c) {
    const float f2 = f * f;
    // sin: f*(s1 + f2*(s3 + f2*(s5 + f2*(s7 + f2*(s9 + f2*(s11 + f2*s13))))))
    const float s1 = 1.5707963267948966f, s3 = -0.6459640975062462f, s5 = 0.07969262624616703f,
                s7 = -0.004681754135318687f, s9 = 0.00016044118478735982f,
                s11 = -3.598843235212085e-06f, s13 = 5.692172921967926e-08f;
    const float c0 = 1.0f, c2 = -1.2337005501361697f, c4 = 0.2536695079010480f,
                c6 = -0.020863480763352957f, c8 = 0.0009192602748394265f,
                c10 = -2.5202042373060604e-05f, c12 = 4.710874778818171e-07f,
                c14 = -6.386603083791069e-09f;
    float ps = s13;
    ps = ps * f2; ps = ps + s11;
    ps = ps * f2; ps = ps + s9;
    ps = ps * f2; ps = ps + s7;
    ps = ps * f2; ps = ps + s5;
    ps = ps * f2; ps = ps + s3;
    ps = ps * f2; ps = ps + s1;
    s = ps * f;
    float pc = c14;
    pc = pc * f2; pc = pc + c12;
    pc = pc * f2; pc = pc + c10;
    pc = pc * f2; pc = pc + c8;
    pc = pc * f2; pc = pc + c6;
    pc = pc * f2; pc = pc + c4;
    pc = pc * f2; pc = pc + c2;
    pc = pc * f2; pc = pc + c0;
    c = pc;
}

// sin and cos of 2*pi*w for w in [0, 1).
inline void sincos_turns(float w, float& s, float& c) {
    float w4 = w * 4.0f;
    float q = std::floor(w4);
    float f = w4 - q;
    float qs, qc;
    sincos_quarter(f, qs, qc);
    int qi = (int)q & 3;
    if (qi == 0)      { s = qs;  c = qc; }
    else if (qi == 1) { s = qc;  c = -qs; }
    else if (qi == 2) { s = -qs; c = -qc; }
    else              { s = -qc; c = qs; }
}

// light.frag:65 uses theta = 6.283 * v (not 2*pi): in turns, v * 6.283/(2*pi).
constexpr float kHemiTurns = 0.9999705074463785f;

// ---------------------------------------------------------------- S3
struct TriRay {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

inline TriRay tri_setup(f3 d) {
    TriRay r;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (comp(d, kz) < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sz = 1.0f / comp(d, kz);
    r.Sx = comp(d, kx) * r.Sz;
    r.Sy = comp(d, ky) * r.Sz;
    return r;
}

// a*b - c*d to within 2 ulps by Kahan's algorithm with fused multiply-adds
// (Jeannerod, Louvet and Muller, Math. Comp. 82, 2013): e = w - c*d exactly,
// so the result is 0 only for an exact 0 and otherwise carries the exact
// difference's sign.  It replaces the published FP64 re-evaluation of zero
// edge functions (Woop, Benthin and Wald 2013, section 3.3).
inline float diff_of_products(float a, float b, float c, float d) {
    const float w = c * d;
    const float e = std::fma(-c, d, w);
    const float f = std::fma(a, b, -w);
    return f + e;
}

// Returns true on hit with T/det in [tmin, tmax]; outputs T, det, U, V, W (signed).
inline bool tri_test(const TriRay& r, f3 o, float tmin, float tmax, f3 v0, f3 v1, f3 v2,
                     float& T_out, float& det_out, float& U_out, float& V_out, float& W_out) {
    f3 A = sub(v0, o), B = sub(v1, o), C = sub(v2, o);
    float Ax = comp(A, r.kx) - r.Sx * comp(A, r.kz);
    float Ay = comp(A, r.ky) - r.Sy * comp(A, r.kz);
    float Bx = comp(B, r.kx) - r.Sx * comp(B, r.kz);
    float By = comp(B, r.ky) - r.Sy * comp(B, r.kz);
    float Cx = comp(C, r.kx) - r.Sx * comp(C, r.kz);
    float Cy = comp(C, r.ky) - r.Sy * comp(C, r.kz);
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
    float Az = r.Sz * comp(A, r.kz);
    float Bz = r.Sz * comp(B, r.kz);
    float Cz = r.Sz * comp(C, r.kz);
    float T = U * Az + V * Bz + W * Cz;
    float adet = det < 0.0f ? -det : det;
    float aT = det < 0.0f ? -T : T;
    if (aT < tmin * adet || aT > tmax * adet) return false;
    T_out = T; det_out = det; U_out = U; V_out = V; W_out = W;
    return true;
}

// FP64 variant of the same test (no fallback needed).
inline bool tri_test_d(d3 o, d3 d, double tmin, double tmax, d3 v0, d3 v1, d3 v2, double& t_out) {
    double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = 0; double m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (dcomp(d, kz) < 0.0) std::swap(kx, ky);
    double Sx = dcomp(d, kx) / dcomp(d, kz), Sy = dcomp(d, ky) / dcomp(d, kz), Sz = 1.0 / dcomp(d, kz);
    d3 A = dsub(v0, o), B = dsub(v1, o), C = dsub(v2, o);
    double Ax = dcomp(A, kx) - Sx * dcomp(A, kz), Ay = dcomp(A, ky) - Sy * dcomp(A, kz);
    double Bx = dcomp(B, kx) - Sx * dcomp(B, kz), By = dcomp(B, ky) - Sy * dcomp(B, kz);
    double Cx = dcomp(C, kx) - Sx * dcomp(C, kz), Cy = dcomp(C, ky) - Sy * dcomp(C, kz);
    double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    double det = U + V + W;
    if (det == 0.0) return false;
    double T = U * Sz * dcomp(A, kz) + V * Sz * dcomp(B, kz) + W * Sz * dcomp(C, kz);
    double t = T / det;
    if (t < tmin || t > tmax) return false;
    t_out = t;
    return true;
}

// ---------------------------------------------------------------- S4
struct BoxRay {
    f3 o, inv;
    float tlo, thi;
};

inline float t_margin(float tmax) {
    float a = std::fabs(tmax);
    return (a > 1.0f ? a : 1.0f) * 0x1p-10f;
}

inline BoxRay box_setup(f3 o, f3 d, float tmin, float tmax) {
    BoxRay b;
    b.o = o;
    b.inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float m = t_margin(tmax);
    b.tlo = tmin - m;
    b.thi = tmax + m;
    return b;
}

inline bool box_hit(const BoxRay& r, float lox, float hix, float loy, float hiy, float loz,
                    float hiz, float& tnear) {
    float nx = r.inv.x < 0.0f ? hix : lox, fx = r.inv.x < 0.0f ? lox : hix;
    float ny = r.inv.y < 0.0f ? hiy : loy, fy = r.inv.y < 0.0f ? loy : hiy;
    float nz = r.inv.z < 0.0f ? hiz : loz, fz = r.inv.z < 0.0f ? loz : hiz;
    float tnx = (nx - r.o.x) * r.inv.x, tfx = (fx - r.o.x) * r.inv.x;
    float tny = (ny - r.o.y) * r.inv.y, tfy = (fy - r.o.y) * r.inv.y;
    float tnz = (nz - r.o.z) * r.inv.z, tfz = (fz - r.o.z) * r.inv.z;
    float tn = r.tlo;
    tn = tnx > tn ? tnx : tn;
    tn = tny > tn ? tny : tn;
    tn = tnz > tn ? tnz : tn;
    float tf = r.thi;
    tf = tfx < tf ? tfx : tf;
    tf = tfy < tf ? tfy : tf;
    tf = tfz < tf ? tfz : tf;
    tf = tf + std::fabs(tf) * 0x1p-20f;
    tnear = tn;
    return tn <= tf;
}

struct DBoxRay { d3 o, inv; double tlo, thi; };
inline bool box_hit_d(const DBoxRay& r, const float* lo, const float* hi) {
    double tn = r.tlo, tf = r.thi;
    const double o[3] = {r.o.x, r.o.y, r.o.z}, inv[3] = {r.inv.x, r.inv.y, r.inv.z};
    for (int a = 0; a < 3; ++a) {
        double n = inv[a] < 0 ? hi[a] : lo[a], f = inv[a] < 0 ? lo[a] : hi[a];
        double t0 = (n - o[a]) * inv[a], t1 = (f - o[a]) * inv[a];
        tn = t0 > tn ? t0 : tn;
        tf = t1 < tf ? t1 : tf;
    }
    tf = tf + std::fabs(tf) * 0x1p-30;
    return tn <= tf;
}

// ---------------------------------------------------------------- S5
struct Node {               // 64 bytes; identical layout on the device
    float lox[2], hix[2];
    float loy[2], hiy[2];
    float loz[2], hiz[2];
    uint32_t child[2];
    uint32_t parent;
    uint32_t pad;
};
static_assert(sizeof(Node) == 64, "node layout");

struct Box { float lo[3], hi[3]; };

inline Box box_union(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = fmin_(a.lo[k], b.lo[k]); r.hi[k] = fmax_(a.hi[k], b.hi[k]); }
    return r;
}

inline Box empty_box() {
    Box b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::numeric_limits<float>::infinity();
        b.hi[k] = -std::numeric_limits<float>::infinity();
    }
    return b;
}

inline void set_child(Node& n, int side, const Box& b) {
    n.lox[side] = b.lo[0]; n.hix[side] = b.hi[0];
    n.loy[side] = b.lo[1]; n.hiy[side] = b.hi[1];
    n.loz[side] = b.lo[2]; n.hiz[side] = b.hi[2];
}

inline Box get_child(const Node& n, int side) {
    Box b;
    b.lo[0] = n.lox[side]; b.hi[0] = n.hix[side];
    b.lo[1] = n.loy[side]; b.hi[1] = n.hiy[side];
    b.lo[2] = n.loz[side]; b.hi[2] = n.hiz[side];
    return b;
}

inline Box tri_box(f3 a, f3 b, f3 c) {
    Box r;
    r.lo[0] = fmin_(fmin_(a.x, b.x), c.x); r.hi[0] = fmax_(fmax_(a.x, b.x), c.x);
    r.lo[1] = fmin_(fmin_(a.y, b.y), c.y); r.hi[1] = fmax_(fmax_(a.y, b.y), c.y);
    r.lo[2] = fmin_(fmin_(a.z, b.z), c.z); r.hi[2] = fmax_(fmax_(a.z, b.z), c.z);
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) {
        m = fmax_(m, std::fabs(r.lo[k]));
        m = fmax_(m, std::fabs(r.hi[k]));
    }
    float pad = m * 0x1p-16f;
    for (int k = 0; k < 3; ++k) { r.lo[k] = r.lo[k] - pad; r.hi[k] = r.hi[k] + pad; }
    return r;
}

inline uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// Karras 2012 delta with index tie-break for equal codes.
inline int delta(const std::vector<uint32_t>& k, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

// The one-node tree of zero or one primitive.
void build_single(const std::vector<Box>& boxes, std::vector<Node>& nodes, std::vector<uint32_t>& order) {
    const int n = (int)boxes.size();
    Node root;
    std::memset(&root, 0, sizeof(root));
    set_child(root, 0, n == 1 ? boxes[0] : empty_box());
    set_child(root, 1, empty_box());
    root.child[0] = n == 1 ? (kLeaf | 0u) : kEmpty;
    root.child[1] = kEmpty;
    root.parent = kNone;
    nodes.push_back(root);
    if (n == 1) order.push_back(0);
}

// Slot order of n >= 2 primitive boxes (S5): stable sort by the Morton code of
// the FP32 box centroid in the centroid frame, whose bounds are reduced in
// FP64 (exact centroids) and then rounded to an FP32 origin and scale.
// order[slot] = primitive index, keys[slot] = its code.
void morton_order(const std::vector<Box>& boxes, std::vector<uint32_t>& order, std::vector<uint32_t>& keys) {
    const int n = (int)boxes.size();
    // FP64 centroid frame.
    std::vector<double> cx(n), cy(n), cz(n);
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n; ++i) {
        cx[i] = 0.5 * ((double)boxes[i].lo[0] + (double)boxes[i].hi[0]);
        cy[i] = 0.5 * ((double)boxes[i].lo[1] + (double)boxes[i].hi[1]);
        cz[i] = 0.5 * ((double)boxes[i].lo[2] + (double)boxes[i].hi[2]);
        double c[3] = {cx[i], cy[i], cz[i]};
        for (int k = 0; k < 3; ++k) {
            mn[k] = c[k] < mn[k] ? c[k] : mn[k];
            mx[k] = c[k] > mx[k] ? c[k] : mx[k];
        }
    }
    float org[3], sc[3];
    for (int k = 0; k < 3; ++k) {
        double e = mx[k] - mn[k];
        org[k] = (float)mn[k];
        sc[k] = e > 0.0 ? (float)(1024.0 / e) : 0.0f;
    }
    std::vector<uint32_t> code(n);
    for (int i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float c = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
            float t = std::floor((c - org[k]) * sc[k]);
            if (!(t >= 0.0f)) t = 0.0f;
            if (t > 1023.0f) t = 1023.0f;
            q[k] = (uint32_t)t;
        }
        code[i] = morton3(q[0], q[1], q[2]);
    }
    order.resize(n);
    for (int i = 0; i < n; ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    keys.resize(n);
    for (int i = 0; i < n; ++i) keys[i] = code[order[i]];
}

// Builds nodes over primitive boxes (S5).  order[slot] = primitive index.
void build_lbvh(const std::vector<Box>& boxes, std::vector<Node>& nodes, std::vector<uint32_t>& order) {
    const int n = (int)boxes.size();
    nodes.clear();
    order.clear();
    if (n <= 1) return build_single(boxes, nodes, order);
    std::vector<uint32_t> keys;
    morton_order(boxes, order, keys);
    nodes.assign(n - 1, Node{});
    for (auto& nd : nodes) std::memset(&nd, 0, sizeof(Node));
    std::vector<uint32_t> leaf_parent(n, kNone);
    nodes[0].parent = kNone;
    for (int i = 0; i < n - 1; ++i) {
        int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        int j = i + l * d;
        int dnode = delta(keys, n, i, j);
        int s = 0;
        int t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int gamma = i + s * d + (d < 0 ? d : 0);
        int lo = i < j ? i : j, hi = i < j ? j : i;
        uint32_t left = (lo == gamma) ? (kLeaf | (uint32_t)gamma) : (uint32_t)gamma;
        uint32_t right = (hi == gamma + 1) ? (kLeaf | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        nodes[i].child[0] = left;
        nodes[i].child[1] = right;
        if (left & kLeaf) leaf_parent[left & ~kLeaf] = (uint32_t)i; else nodes[left].parent = (uint32_t)i;
        if (right & kLeaf) leaf_parent[right & ~kLeaf] = (uint32_t)i; else nodes[right].parent = (uint32_t)i;
    }
    // Refit: post-order from the root (exact min/max, order independent).
    std::vector<Box> nb(n - 1);
    std::vector<std::pair<uint32_t, int>> st;
    st.push_back({0u, 0});
    while (!st.empty()) {
        auto& top = st.back();
        uint32_t ni = top.first;
        if (top.second < 2) {
            int side = top.second++;
            uint32_t c = nodes[ni].child[side];
            if (!(c & kLeaf)) st.push_back({c, 0});
            continue;
        }
        Box cb[2];
        for (int side = 0; side < 2; ++side) {
            uint32_t c = nodes[ni].child[side];
            cb[side] = (c & kLeaf) ? boxes[order[c & ~kLeaf]] : nb[c];
            set_child(nodes[ni], side, cb[side]);
        }
        nb[ni] = box_union(cb[0], cb[1]);
        st.pop_back();
    }
}

// S7: TLAS hierarchy by locally-ordered clustering (PLOC, Meister & Bittner
// 2018, IEEE TVCG 24(3)) over the Morton slot order.  Clusters start as the
// leaves in slot order.  Each round, cluster i takes as nearest neighbour the
// j within kPlocRadius positions minimising the merge cost (half the surface
// area of the union box, FP32, evaluated as written; NaN counts as +inf),
// the smaller j on equal cost.  Mutual nearest neighbours i < j merge into a
// new node (child 0 = cluster i, child 1 = cluster j) that takes position i;
// position j drops out; the survivors keep their relative order.  Nodes are
// numbered downwards from n - 2 in the order they are made (merges of one
// round in increasing i), so the last merge, the root, is node 0.  The pair
// of least (cost, i, j) is always mutual, so every round merges.
constexpr int kPlocRadius = 16;

inline float merge_cost(const Box& a, const Box& b) {
    const Box u = box_union(a, b);
    const float ex = u.hi[0] - u.lo[0], ey = u.hi[1] - u.lo[1], ez = u.hi[2] - u.lo[2];
    const float c = ex * ey + ey * ez + ez * ex;
    return c == c ? c : INFINITY;
}

void build_ploc(const std::vector<Box>& boxes, std::vector<Node>& nodes, std::vector<uint32_t>& order) {
    const int n = (int)boxes.size();
    nodes.clear();
    order.clear();
    if (n <= 1) return build_single(boxes, nodes, order);
    std::vector<uint32_t> keys;
    morton_order(boxes, order, keys);
    nodes.assign(n - 1, Node{});
    for (auto& nd : nodes) std::memset(&nd, 0, sizeof(Node));
    std::vector<uint32_t> ref(n), nn(n), ref2(n);
    std::vector<Box> cb(n), cb2(n);
    for (int s = 0; s < n; ++s) {
        ref[s] = kLeaf | (uint32_t)s;
        cb[s] = boxes[order[s]];
    }
    int count = n;
    uint32_t next = (uint32_t)n - 1;
    while (count > 1) {
        for (int i = 0; i < count; ++i) {
            float best = INFINITY;
            uint32_t bj = kNone;
            const int j0 = i - kPlocRadius < 0 ? 0 : i - kPlocRadius;
            const int j1 = i + kPlocRadius > count - 1 ? count - 1 : i + kPlocRadius;
            for (int j = j0; j <= j1; ++j) {
                if (j == i) continue;
                const float c = merge_cost(cb[i], cb[j]);
                if (bj == kNone || c < best) { best = c; bj = (uint32_t)j; }
            }
            nn[i] = bj;
        }
        int w = 0;
        uint32_t made = 0;
        for (int i = 0; i < count; ++i) {
            const uint32_t j = nn[i];
            const bool mutual = nn[j] == (uint32_t)i;
            if (mutual && (uint32_t)i > j) continue;
            if (mutual) {
                const uint32_t idx = next - 1 - made++;
                Node& nd = nodes[idx];
                nd.child[0] = ref[i];
                nd.child[1] = ref[j];
                set_child(nd, 0, cb[i]);
                set_child(nd, 1, cb[j]);
                nd.parent = kNone;
                for (int side = 0; side < 2; ++side)
                    if (!(nd.child[side] & kLeaf)) nodes[nd.child[side]].parent = idx;
                ref2[w] = idx;
                cb2[w] = box_union(cb[i], cb[j]);
            } else {
                ref2[w] = ref[i];
                cb2[w] = cb[i];
            }
            ++w;
        }
        count = w;
        next -= made;
        std::swap(ref, ref2);
        std::swap(cb, cb2);
    }
}

inline Box root_box(const Node& r) { return box_union(get_child(r, 0), get_child(r, 1)); }

// ---------------------------------------------------------------- S7
// 8-wide BVH with 8-bit child boxes (the precision form).  Collapse of the
// binary LBVH: a wide node starts from a binary node's two children and
// repeatedly replaces its internal child of largest surface area (first on
// ties) by that child's two children, up to 8 slots.  Wide nodes are numbered
// breadth first; internal slots point to child_base + rank, leaf slots to
// triangle tri_base + rank (leaves hold one triangle).  Child boxes are
// quantised per axis in a power-of-two frame: scale = 2^e, origin a multiple
// of 2^e, lo rounded down and hi rounded up, so that decode (origin + q*2^e,
// exact in FP32) always contains the child box.
struct Node8 {              // 96 bytes; identical layout on the device
    float ox, oy, oz;
    int8_t ex, ey, ez;
    uint8_t nchild;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];        // bit 7: internal child; bits 0-6: rank among internal or leaf slots
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
    uint32_t pad[4];
};
static_assert(sizeof(Node8) == 96, "wide node layout");

struct Ref { uint32_t c; Box b; };

inline float surface_area(const Box& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

void wide_collapse(const std::vector<Node>& bn, uint32_t node, std::vector<Ref>& slots) {
    slots.clear();
    for (int c = 0; c < 2; ++c)
        if (bn[node].child[c] != kEmpty) slots.push_back(Ref{bn[node].child[c], get_child(bn[node], c)});
    while (slots.size() < 8) {
        int best = -1;
        float best_sa = 0.0f;
        for (size_t k = 0; k < slots.size(); ++k) {
            if (slots[k].c & kLeaf) continue;
            float sa = surface_area(slots[k].b);
            if (best < 0 || sa > best_sa) { best = (int)k; best_sa = sa; }
        }
        if (best < 0) break;
        uint32_t in = slots[best].c;
        Ref a{bn[in].child[0], get_child(bn[in], 0)}, b{bn[in].child[1], get_child(bn[in], 1)};
        slots[best] = a;
        slots.insert(slots.begin() + best + 1, b);
    }
}

void quant_frame(float lo, float hi, float& origin, int& e) {
    float t = (hi - lo) / 255.0f;
    int e0 = -126, em = -126;
    if (t > 0.0f) std::frexp(t, &e0);
    float mag = std::fabs(lo) > std::fabs(hi) ? std::fabs(lo) : std::fabs(hi);
    if (mag > 0.0f) std::frexp(mag, &em);
    e = e0 > em - 23 ? e0 : em - 23;
    if (e < -126) e = -126;
    for (;;) {
        float sc = std::ldexp(1.0f, e);
        origin = std::floor(lo / sc) * sc;
        if (origin > lo) origin = origin - sc;
        if (origin + 255.0f * sc >= hi) break;
        ++e;
    }
}

inline uint8_t quant_lo(float v, float origin, float sc) {
    float q = std::floor((v - origin) / sc);
    q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
    while (q > 0.0f && origin + q * sc > v) q = q - 1.0f;
    return (uint8_t)q;
}

inline uint8_t quant_hi(float v, float origin, float sc) {
    float q = std::ceil((v - origin) / sc);
    q = q < 0.0f ? 0.0f : (q > 255.0f ? 255.0f : q);
    while (q < 255.0f && origin + q * sc < v) q = q + 1.0f;
    return (uint8_t)q;
}

inline float wide_scale(int8_t e) { return std::ldexp(1.0f, (int)e); }

// Decoded box of slot k (exact FP32).
inline void wide_box(const Node8& n, int k, float lo[3], float hi[3]) {
    const float o[3] = {n.ox, n.oy, n.oz};
    const int8_t ex[3] = {n.ex, n.ey, n.ez};
    for (int a = 0; a < 3; ++a) {
        float sc = wide_scale(ex[a]);
        lo[a] = o[a] + (float)n.qlo[a][k] * sc;
        hi[a] = o[a] + (float)n.qhi[a][k] * sc;
    }
}

void build_wide(const std::vector<Node>& bn, const std::vector<float>& tris, const std::vector<uint32_t>& prims,
                std::vector<Node8>& out, std::vector<float>& tris8, std::vector<uint32_t>& prims8) {
    out.clear();
    tris8.clear();
    prims8.clear();
    std::vector<uint32_t> level = {0u};
    uint32_t level_start = 0;
    std::vector<Ref> slots;
    while (!level.empty()) {
        const uint32_t next_start = level_start + (uint32_t)level.size();
        std::vector<uint32_t> next;
        for (uint32_t node : level) {
            wide_collapse(bn, node, slots);
            Node8 w;
            std::memset(&w, 0, sizeof(w));
            w.nchild = (uint8_t)slots.size();
            w.child_base = next_start + (uint32_t)next.size();
            w.tri_base = (uint32_t)prims8.size();
            Box nb = empty_box();
            for (auto& r : slots) nb = box_union(nb, r.b);
            float origin[3];
            int e[3] = {0, 0, 0};
            for (int a = 0; a < 3; ++a) {
                if (slots.empty()) { origin[a] = 0.0f; e[a] = 0; continue; }
                quant_frame(nb.lo[a], nb.hi[a], origin[a], e[a]);
            }
            w.ox = origin[0]; w.oy = origin[1]; w.oz = origin[2];
            w.ex = (int8_t)e[0]; w.ey = (int8_t)e[1]; w.ez = (int8_t)e[2];
            uint32_t ni = 0, nl = 0;
            for (size_t k = 0; k < slots.size(); ++k) {
                for (int a = 0; a < 3; ++a) {
                    float sc = std::ldexp(1.0f, e[a]);
                    w.qlo[a][k] = quant_lo(slots[k].b.lo[a], origin[a], sc);
                    w.qhi[a][k] = quant_hi(slots[k].b.hi[a], origin[a], sc);
                }
                if (slots[k].c & kLeaf) {
                    uint32_t slot = slots[k].c & ~kLeaf;
                    w.meta[k] = (uint8_t)nl++;
                    tris8.insert(tris8.end(), tris.begin() + (size_t)slot * 9, tris.begin() + (size_t)slot * 9 + 9);
                    prims8.push_back(prims[slot]);
                } else {
                    w.meta[k] = (uint8_t)(0x80u | ni++);
                    next.push_back(slots[k].c);
                }
            }
            out.push_back(w);
        }
        level_start = next_start;
        level.swap(next);
    }
}

// ---------------------------------------------------------------- S6
void invert_3x4(const float m[12], float out[12], double* dout = nullptr) {
    double r00 = m[0], r01 = m[1], r02 = m[2], t0 = m[3];
    double r10 = m[4], r11 = m[5], r12 = m[6], t1 = m[7];
    double r20 = m[8], r21 = m[9], r22 = m[10], t2 = m[11];
    double c00 = r11 * r22 - r12 * r21;
    double c01 = r12 * r20 - r10 * r22;
    double c02 = r10 * r21 - r11 * r20;
    double det = r00 * c00 + r01 * c01 + r02 * c02;
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
        for (int j = 0; j < 3; ++j) {
            out[i * 4 + j] = (float)inv[i][j];
            if (dout) dout[i * 4 + j] = inv[i][j];
        }
        out[i * 4 + 3] = (float)ti;
        if (dout) dout[i * 4 + 3] = ti;
    }
}

inline f3 xform_point(const float* w, f3 p) {
    return mk(w[0] * p.x + w[1] * p.y + w[2] * p.z + w[3],
              w[4] * p.x + w[5] * p.y + w[6] * p.z + w[7],
              w[8] * p.x + w[9] * p.y + w[10] * p.z + w[11]);
}
inline f3 xform_vec(const float* w, f3 v) {
    return mk(w[0] * v.x + w[1] * v.y + w[2] * v.z,
              w[4] * v.x + w[5] * v.y + w[6] * v.z,
              w[8] * v.x + w[9] * v.y + w[10] * v.z);
}

struct Blas {
    std::vector<Node> nodes;
    std::vector<float> tris;      // 9 floats per slot
    std::vector<float> nrms;      // 9 floats per slot
    std::vector<uint32_t> prims;  // slot -> primitive
    std::vector<uint32_t> indices;
    uint32_t vertex_count = 0;
    uint32_t stride = 48;
    bool has_normals = false;
    // precision form
    std::vector<Node8> wide;
    std::vector<float> tris8;
    std::vector<uint32_t> prims8;
};

struct Inst {
    float o2w[12];
    float w2o[12];
    double w2o_d[12];
    float nmat[9];   // inverse-transpose of the 3x3 part (row-major)
    uint32_t mesh, custom, mask, input_index;
};

}  // namespace

struct orc_scene {
    std::vector<Blas> blas;
    std::vector<Inst> inst_input;  // all instances, input order
    std::vector<Node> tlas_nodes;
    std::vector<uint32_t> tlas_order;  // slot -> index into tlas_inst
    std::vector<Inst> tlas_inst;       // instances kept in the TLAS (input order)
    std::vector<float> tlas_w2o;       // 12 floats per slot
};

namespace {

void blas_build(Blas& b, const void* vertices, uint32_t stride, uint32_t nv) {
    const uint8_t* base = (const uint8_t*)vertices;
    const uint32_t nt = (uint32_t)(b.indices.size() / 3);
    std::vector<Box> boxes(nt);
    std::vector<f3> pos(nv), nrm(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        const float* p = (const float*)(base + (size_t)i * stride);
        pos[i] = mk(p[0], p[1], p[2]);
        nrm[i] = b.has_normals ? mk(p[3], p[4], p[5]) : mk(0, 0, 0);
    }
    for (uint32_t t = 0; t < nt; ++t)
        boxes[t] = tri_box(pos[b.indices[3 * t]], pos[b.indices[3 * t + 1]], pos[b.indices[3 * t + 2]]);
    build_lbvh(boxes, b.nodes, b.prims);
    b.tris.assign((size_t)nt * 9, 0.0f);
    b.nrms.assign((size_t)nt * 9, 0.0f);
    for (uint32_t s = 0; s < nt; ++s) {
        uint32_t t = b.prims[s];
        for (int v = 0; v < 3; ++v) {
            f3 p = pos[b.indices[3 * t + v]], q = nrm[b.indices[3 * t + v]];
            b.tris[(size_t)s * 9 + v * 3 + 0] = p.x;
            b.tris[(size_t)s * 9 + v * 3 + 1] = p.y;
            b.tris[(size_t)s * 9 + v * 3 + 2] = p.z;
            b.nrms[(size_t)s * 9 + v * 3 + 0] = q.x;
            b.nrms[(size_t)s * 9 + v * 3 + 1] = q.y;
            b.nrms[(size_t)s * 9 + v * 3 + 2] = q.z;
        }
    }
}

// Refit keeps the topology and primitive order; recomputes boxes bottom-up.
void blas_refit(Blas& b, const void* vertices, uint32_t stride) {
    const uint8_t* base = (const uint8_t*)vertices;
    const uint32_t nt = (uint32_t)(b.indices.size() / 3);
    auto P = [&](uint32_t i) { const float* p = (const float*)(base + (size_t)i * stride); return mk(p[0], p[1], p[2]); };
    std::vector<Box> slot_box(nt);
    for (uint32_t s = 0; s < nt; ++s) {
        uint32_t t = b.prims[s];
        f3 a = P(b.indices[3 * t]), bb = P(b.indices[3 * t + 1]), c = P(b.indices[3 * t + 2]);
        slot_box[s] = tri_box(a, bb, c);
        const f3 vv[3] = {a, bb, c};
        for (int v = 0; v < 3; ++v) {
            b.tris[(size_t)s * 9 + v * 3 + 0] = vv[v].x;
            b.tris[(size_t)s * 9 + v * 3 + 1] = vv[v].y;
            b.tris[(size_t)s * 9 + v * 3 + 2] = vv[v].z;
        }
    }
    if (nt <= 1) {
        set_child(b.nodes[0], 0, nt == 1 ? slot_box[0] : empty_box());
        return;
    }
    std::vector<Box> nb(b.nodes.size());
    std::vector<std::pair<uint32_t, int>> st;
    st.push_back({0u, 0});
    while (!st.empty()) {
        auto& top = st.back();
        uint32_t ni = top.first;
        if (top.second < 2) {
            int side = top.second++;
            uint32_t c = b.nodes[ni].child[side];
            if (!(c & kLeaf)) st.push_back({c, 0});
            continue;
        }
        Box cb[2];
        for (int side = 0; side < 2; ++side) {
            uint32_t c = b.nodes[ni].child[side];
            cb[side] = (c & kLeaf) ? slot_box[c & ~kLeaf] : nb[c];
            set_child(b.nodes[ni], side, cb[side]);
        }
        nb[ni] = box_union(cb[0], cb[1]);
        st.pop_back();
    }
}

void tlas_build(orc_scene& s) {
    s.tlas_inst.clear();
    for (const Inst& in : s.inst_input) {
        const Blas& b = s.blas[in.mesh];
        // Empty meshes and mask-0 instances can never be hit by a 0xFF ray mask.
        if (b.prims.empty() || (in.mask & 0xFFu) == 0) continue;
        s.tlas_inst.push_back(in);
    }
    const uint32_t n = (uint32_t)s.tlas_inst.size();
    std::vector<Box> boxes(n);
    float scene_m = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const Inst& in = s.tlas_inst[i];
        Box rb = root_box(s.blas[in.mesh].nodes[0]);
        Box wb;
        for (int k = 0; k < 3; ++k) { wb.lo[k] = INFINITY; wb.hi[k] = -INFINITY; }
        for (int c = 0; c < 8; ++c) {
            f3 p = mk((c & 1) ? rb.hi[0] : rb.lo[0], (c & 2) ? rb.hi[1] : rb.lo[1], (c & 4) ? rb.hi[2] : rb.lo[2]);
            f3 w = xform_point(in.o2w, p);
            float wc[3] = {w.x, w.y, w.z};
            for (int k = 0; k < 3; ++k) { wb.lo[k] = fmin_(wb.lo[k], wc[k]); wb.hi[k] = fmax_(wb.hi[k], wc[k]); }
        }
        for (int k = 0; k < 3; ++k) {
            scene_m = fmax_(scene_m, std::fabs(wb.lo[k]));
            scene_m = fmax_(scene_m, std::fabs(wb.hi[k]));
        }
        boxes[i] = wb;
    }
    float pad = scene_m * 0x1p-14f;
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) { boxes[i].lo[k] = boxes[i].lo[k] - pad; boxes[i].hi[k] = boxes[i].hi[k] + pad; }
    build_ploc(boxes, s.tlas_nodes, s.tlas_order);
    s.tlas_w2o.assign((size_t)n * 12, 0.0f);
    for (uint32_t slot = 0; slot < n; ++slot)
        std::memcpy(&s.tlas_w2o[(size_t)slot * 12], s.tlas_inst[s.tlas_order[slot]].w2o, 48);
}

// ------------------------------------------------------------ traversal
struct Hit {
    float t;
    uint32_t inst, slot;   // instance input index, BLAS slot
    float T, det, U, V, W;
};

struct LevelRay {
    f3 o, d;
    TriRay tr;
    BoxRay br;
};

inline LevelRay level_ray(f3 o, f3 d, float tmin, float tmax) {
    LevelRay r;
    r.o = o; r.d = d;
    r.tr = tri_setup(d);
    r.br = box_setup(o, d, tmin, tmax);
    return r;
}

// Canonical any-hit (closest == false) or closest-hit traversal.
bool traverse(const orc_scene& s, f3 wo, f3 wd, float tmin, float tmax, bool closest, Hit* best,
              orc_stats* st) {
    if (s.tlas_inst.empty()) return false;
    LevelRay world = level_ray(wo, wd, tmin, tmax);
    LevelRay obj;
    const Blas* blas = nullptr;
    bool in_blas = false;
    uint32_t cur_inst = 0;
    float best_t = tmax;
    bool found = false;
    uint32_t stack[256];
    int sp = 0;
    uint32_t next = 0;  // TLAS root
    for (;;) {
        if (next == kNone) {
            if (sp == 0) break;
            next = stack[--sp];
            if (next == kSentinel) {
                in_blas = false;
                next = kNone;
                continue;
            }
        }
        if (!in_blas && (next & kLeaf)) {
            // enter instance
            uint32_t slot = next & ~kLeaf;
            const Inst& in = s.tlas_inst[s.tlas_order[slot]];
            if (st) st->inst_visits++;
            const float* w = &s.tlas_w2o[(size_t)slot * 12];
            obj = level_ray(xform_point(w, wo), xform_vec(w, wd), tmin, best_t);
            blas = &s.blas[in.mesh];
            cur_inst = in.input_index;
            in_blas = true;
            stack[sp++] = kSentinel;
            next = 0;
            continue;
        }
        const LevelRay& r = in_blas ? obj : world;
        const Node& nd = in_blas ? blas->nodes[next] : s.tlas_nodes[next];
        if (st) st->node_visits++;
        float tn[2];
        bool h[2];
        for (int c = 0; c < 2; ++c)
            h[c] = nd.child[c] != kEmpty &&
                   box_hit(r.br, nd.lox[c], nd.hix[c], nd.loy[c], nd.hiy[c], nd.loz[c], nd.hiz[c], tn[c]);
        int first = (h[0] && h[1]) ? (tn[1] < tn[0] ? 1 : 0) : (h[0] ? 0 : 1);
        next = kNone;
        for (int k = 0; k < 2; ++k) {
            int c = k == 0 ? first : 1 - first;
            if (!h[c]) continue;
            uint32_t ch = nd.child[c];
            if (in_blas && (ch & kLeaf)) {
                uint32_t slot = ch & ~kLeaf;
                const float* v = &blas->tris[(size_t)slot * 9];
                if (st) st->tri_tests++;
                float T, det, U, V, W;
                if (tri_test(r.tr, r.o, tmin, tmax, load3(v), load3(v + 3), load3(v + 6), T, det, U, V, W)) {
                    if (!closest) {
                        if (st) st->hits++;
                        return true;
                    }
                    float t = T / det;
                    bool better = !found || t < best->t ||
                                  (t == best->t && (cur_inst < best->inst ||
                                                    (cur_inst == best->inst && blas->prims[slot] < blas->prims[best->slot])));
                    if (better) {
                        found = true;
                        best->t = t; best->inst = cur_inst; best->slot = slot;
                        best->T = T; best->det = det; best->U = U; best->V = V; best->W = W;
                        best_t = t;
                        float m = t_margin(t);
                        obj.br.thi = t + m;
                        world.br.thi = t + m;
                    }
                }
                continue;
            }
            if (next == kNone) next = ch;
            else if (sp < 255) stack[sp++] = ch;
            else return closest ? found : true;  // unreachable for LBVH depth <= 127
        }
    }
    if (found && st) st->hits++;
    return found;
}

// Brute force over every triangle of every instance (same ray transform).
bool brute(const orc_scene& s, f3 wo, f3 wd, float tmin, float tmax, bool closest, Hit* best) {
    bool found = false;
    for (uint32_t slot = 0; slot < s.tlas_order.size(); ++slot) {
        const Inst& in = s.tlas_inst[s.tlas_order[slot]];
        const float* w = &s.tlas_w2o[(size_t)slot * 12];
        f3 o = xform_point(w, wo), d = xform_vec(w, wd);
        TriRay tr = tri_setup(d);
        const Blas& b = s.blas[in.mesh];
        for (uint32_t ts = 0; ts < b.prims.size(); ++ts) {
            const float* v = &b.tris[(size_t)ts * 9];
            float T, det, U, V, W;
            if (!tri_test(tr, o, tmin, tmax, load3(v), load3(v + 3), load3(v + 6), T, det, U, V, W)) continue;
            if (!closest) return true;
            float t = T / det;
            bool better = !found || t < best->t ||
                          (t == best->t && (in.input_index < best->inst ||
                                            (in.input_index == best->inst && b.prims[ts] < s.blas[s.inst_input[best->inst].mesh].prims[best->slot])));
            if (better) {
                found = true;
                best->t = t; best->inst = in.input_index; best->slot = ts;
                best->T = T; best->det = det; best->U = U; best->V = V; best->W = W;
            }
        }
    }
    return found;
}

// Any-hit traversal with the 8-wide quantised BLASes (TLAS stays binary).
// Hit slots are visited in slot order: leaves are tested at once, the first
// internal child is visited next and the others are pushed.
bool traverse_wide(const orc_scene& s, f3 wo, f3 wd, float tmin, float tmax, orc_stats* st) {
    if (s.tlas_inst.empty()) return false;
    LevelRay world = level_ray(wo, wd, tmin, tmax);
    LevelRay obj;
    const Blas* blas = nullptr;
    bool in_blas = false;
    uint32_t stack[512];
    int sp = 0;
    uint32_t next = 0;
    for (;;) {
        if (next == kNone) {
            if (sp == 0) break;
            next = stack[--sp];
            if (next == kSentinel) { in_blas = false; next = kNone; continue; }
        }
        if (!in_blas && (next & kLeaf)) {
            uint32_t slot = next & ~kLeaf;
            const Inst& in = s.tlas_inst[s.tlas_order[slot]];
            if (st) st->inst_visits++;
            const float* w = &s.tlas_w2o[(size_t)slot * 12];
            obj = level_ray(xform_point(w, wo), xform_vec(w, wd), tmin, tmax);
            blas = &s.blas[in.mesh];
            in_blas = true;
            stack[sp++] = kSentinel;
            next = 0;
            continue;
        }
        if (!in_blas) {
            const Node& nd = s.tlas_nodes[next];
            if (st) st->node_visits++;
            float tn[2] = {0.0f, 0.0f};
            bool h[2];
            for (int c = 0; c < 2; ++c)
                h[c] = nd.child[c] != kEmpty &&
                       box_hit(world.br, nd.lox[c], nd.hix[c], nd.loy[c], nd.hiy[c], nd.loz[c], nd.hiz[c], tn[c]);
            int first = (h[0] && h[1]) ? (tn[1] < tn[0] ? 1 : 0) : (h[0] ? 0 : 1);
            next = kNone;
            for (int k = 0; k < 2; ++k) {
                int c = k == 0 ? first : 1 - first;
                if (!h[c]) continue;
                if (next == kNone) next = nd.child[c];
                else stack[sp++] = nd.child[c];
            }
            continue;
        }
        const Node8& n = blas->wide[next];
        if (st) st->wide_node_visits++;
        next = kNone;
        for (int k = 0; k < n.nchild; ++k) {
            float lo[3], hi[3], tn;
            wide_box(n, k, lo, hi);
            if (!box_hit(obj.br, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], tn)) continue;
            if (n.meta[k] & 0x80u) {
                uint32_t child = n.child_base + (n.meta[k] & 0x7Fu);
                if (next == kNone) next = child;
                else if (sp < 511) stack[sp++] = child;
                else return true;  // unreachable: depth * 7 < 511
            } else {
                const float* v = &blas->tris8[(size_t)(n.tri_base + n.meta[k]) * 9];
                if (st) st->tri_tests++;
                float T, det, U, V, W;
                if (tri_test(obj.tr, obj.o, tmin, tmax, load3(v), load3(v + 3), load3(v + 6), T, det, U, V, W)) {
                    if (st) st->hits++;
                    return true;
                }
            }
        }
    }
    return false;
}

// FP64 any-hit over the same BVH (boxes tested in FP64 with a margin).
bool traverse_d(const orc_scene& s, d3 wo, d3 wd, double tmin, double tmax) {
    if (s.tlas_inst.empty()) return false;
    auto mkbox = [&](d3 o, d3 d) {
        DBoxRay b;
        b.o = o;
        b.inv = dk(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
        double m = (std::fabs(tmax) > 1.0 ? std::fabs(tmax) : 1.0) * 0x1p-10;
        b.tlo = tmin - m; b.thi = tmax + m;
        return b;
    };
    std::vector<uint32_t> tst;
    tst.push_back(0);
    DBoxRay wb = mkbox(wo, wd);
    while (!tst.empty()) {
        uint32_t ni = tst.back(); tst.pop_back();
        const Node& nd = s.tlas_nodes[ni];
        for (int c = 0; c < 2; ++c) {
            uint32_t ch = nd.child[c];
            if (ch == kEmpty) continue;
            Box cb = get_child(nd, c);
            if (!box_hit_d(wb, cb.lo, cb.hi)) continue;
            if (!(ch & kLeaf)) { tst.push_back(ch); continue; }
            uint32_t slot = ch & ~kLeaf;
            const Inst& in = s.tlas_inst[s.tlas_order[slot]];
            const double* w = in.w2o_d;
            d3 o = dk(w[0] * wo.x + w[1] * wo.y + w[2] * wo.z + w[3],
                      w[4] * wo.x + w[5] * wo.y + w[6] * wo.z + w[7],
                      w[8] * wo.x + w[9] * wo.y + w[10] * wo.z + w[11]);
            d3 d = dk(w[0] * wd.x + w[1] * wd.y + w[2] * wd.z,
                      w[4] * wd.x + w[5] * wd.y + w[6] * wd.z,
                      w[8] * wd.x + w[9] * wd.y + w[10] * wd.z);
            DBoxRay ob = mkbox(o, d);
            const Blas& b = s.blas[in.mesh];
            std::vector<uint32_t> bst;
            bst.push_back(0);
            while (!bst.empty()) {
                uint32_t bi = bst.back(); bst.pop_back();
                const Node& bn = b.nodes[bi];
                for (int k = 0; k < 2; ++k) {
                    uint32_t bc = bn.child[k];
                    if (bc == kEmpty) continue;
                    Box kb = get_child(bn, k);
                    if (!box_hit_d(ob, kb.lo, kb.hi)) continue;
                    if (!(bc & kLeaf)) { bst.push_back(bc); continue; }
                    const float* v = &b.tris[(size_t)(bc & ~kLeaf) * 9];
                    double t;
                    if (tri_test_d(o, d, tmin, tmax, dk(v[0], v[1], v[2]), dk(v[3], v[4], v[5]),
                                   dk(v[6], v[7], v[8]), t))
                        return true;
                }
            }
        }
    }
    return false;
}

// ------------------------------------------------------------ ray set-up
struct PixelRays {
    bool background;
    f3 so;     // shadow origin
    f3 ao_o;   // AO origin
    f3 aoT, aoB, N;
};

inline void unpack8(uint32_t v, float out[4]) {
    for (int k = 0; k < 4; ++k) out[k] = (float)((v >> (8 * k)) & 0xFFu) / 255.0f;
}

// Shadow ray of light l, sample s (light.frag:86-109,137-146).
struct ShadowSetup { f3 Lvec, T, B; float tmax; int n; float radius; };

inline ShadowSetup shadow_setup(const orc_light& L, f3 P) {
    ShadowSetup r;
    f3 lpos = load3(L.position);
    f3 Lraw = sub(lpos, P);
    float dist = length(Lraw);
    f3 Ldir;
    if (L.type == 2) {
        f3 dir = load3(L.direction);
        Ldir = normalize(mk(-dir.x, -dir.y, -dir.z));
        r.Lvec = mul(mul(dir, dot(dir, Ldir)), dist);
    } else {
        Ldir = normalize(Lraw);
        r.Lvec = mul(Ldir, dist);
    }
    f3 c = cross(r.Lvec, mk(0.0f, 1.0f, 0.0f));
    if (length(c) == 0.0f) c = cross(r.Lvec, mk(1.0f, 0.0f, 0.0f));
    r.T = normalize(c);
    r.B = normalize(cross(r.T, r.Lvec));
    r.tmax = length(r.Lvec);
    r.n = L.num_shadow_samples;
    r.radius = L.radius;
    return r;
}

inline f3 shadow_dir(const ShadowSetup& ss, float u, float v) {
    float pr = ss.radius * std::sqrt(u);
    float sn, cs;
    sincos_turns(v, sn, cs);
    float dx = pr * cs, dy = pr * sn;
    f3 d = mk(ss.Lvec.x + dx * ss.T.x + dy * ss.B.x,
              ss.Lvec.y + dx * ss.T.y + dy * ss.B.y,
              ss.Lvec.z + dx * ss.T.z + dy * ss.B.z);
    return normalize(d);
}

inline f3 ao_dir(f3 T, f3 B, f3 N, float u, float v) {
    float r = std::sqrt(u);
    float sn, cs;
    sincos_turns(v * kHemiTurns, sn, cs);
    float hx = r * cs, hy = r * sn;
    float om = 1.0f - u;
    float hz = std::sqrt(om > 0.0f ? om : 0.0f);
    return mk(T.x * hx + B.x * hy + N.x * hz,
              T.y * hx + B.y * hy + N.y * hz,
              T.z * hx + B.z * hy + N.z * hz);
}

inline PixelRays pixel_setup(const orc_gbuffer* g, const orc_params* p, size_t idx) {
    PixelRays r;
    f3 P = load3(g->position + idx * 3), N = load3(g->normal + idx * 3);
    r.N = N;
    r.background = length(N) == 0.0f;
    if (r.background) return r;
    f3 cam = load3(p->cam_pos);
    float lc = length(sub(P, cam));
    float sb = fmax_(lc * 0.01f, 0.05f);
    r.so = add(P, mul(N, sb));
    float ab = lc * 0.01f;
    r.ao_o = add(P, mul(N, ab));
    r.aoT = std::fabs(N.z) > 0.5f ? mk(0.0f, -N.z, N.y) : mk(-N.y, N.x, 0.0f);
    r.aoB = cross(N, r.aoT);
    return r;
}

// Runs fn(y, thread) for rows y = 0, step, 2*step, ... < h.
template <typename F>
void parallel_rows(uint32_t h, int threads, F&& fn, uint32_t step = 1) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (step == 0) step = 1;
    const uint32_t rows = (h + step - 1) / step;
    if (threads <= 1 || rows < 2) {
        for (uint32_t r = 0; r < rows; ++r) fn(r * step, 0);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (;;) {
                uint32_t r = next.fetch_add(1);
                if (r >= rows) break;
                fn(r * step, t);
            }
        });
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------ shading
inline float srgb_to_linear(uint32_t c) { return std::pow((float)c / 255.0f, 2.2f); }

template <typename R>
struct Shade3 { R x, y, z; };

// Shadow rays are cast only in ray-traced scenes: GPUScene.cpp:248 packs
// numShadowSamples = lightSamples only for ShadowRayTraced, and
// EvaluateShadow calls TraceShadowRay only then (light.frag:141).
inline int traced_samples(const orc_params* p, int l) {
    return p->shadow_type == 1 ? p->lights[l].num_shadow_samples : 0;
}

// EvaluateShadow's result by shadow type (light.frag:141-168): ray-traced
// scenes take occluded/samples (0 without samples, TraceShadowRay :87-89);
// shadow-map scenes take the map decision supplied as one sample per light,
// and are fully shadowed for a light without one (light.shadowMap == -1);
// every other type (ShadowDisabled) is fully shadowed.
template <typename R>
inline R shadow_factor(const orc_params* p, const uint8_t* shadow, int l, size_t npix, size_t idx) {
    const int ns = p->lights[l].num_shadow_samples;
    const bool from_counts = p->shadow_type == 1 || (p->shadow_type == 2 && ns > 0);
    if (!from_counts) return (R)1;
    return ns > 0 ? (R)shadow[(size_t)l * npix + idx] / (R)ns : (R)0;
}

template <typename R>
void shade_pixel(const orc_gbuffer* g, const orc_params* p, const uint8_t* shadow,
                 const uint8_t* ao, size_t idx, size_t npix, R out[4]) {
    using std::sqrt; using std::pow;
    auto ld = [](const float* q) { return Shade3<R>{(R)q[0], (R)q[1], (R)q[2]}; };
    auto sub3 = [](Shade3<R> a, Shade3<R> b) { return Shade3<R>{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto add3 = [](Shade3<R> a, Shade3<R> b) { return Shade3<R>{a.x + b.x, a.y + b.y, a.z + b.z}; };
    auto dot3 = [](Shade3<R> a, Shade3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    auto len3 = [&](Shade3<R> a) { return sqrt(dot3(a, a)); };
    auto nrm3 = [&](Shade3<R> a) { R r = (R)1 / len3(a); return Shade3<R>{a.x * r, a.y * r, a.z * r}; };
    auto mx = [](R a, R b) { return a > b ? a : b; };
    auto clamp01 = [](R a) { return a < (R)0 ? (R)0 : (a > (R)1 ? (R)1 : a); };
    const R PI = (R)3.14159265359;

    Shade3<R> N = ld(g->normal + idx * 3);
    if (len3(N) == (R)0) {
        out[0] = (R)p->ambient_color[0] * (R)p->ambient_intensity;
        out[1] = (R)p->ambient_color[1] * (R)p->ambient_intensity;
        out[2] = (R)p->ambient_color[2] * (R)p->ambient_intensity;
        out[3] = (R)1;
        return;
    }
    uint32_t alb = g->albedo[idx], mat = g->material[idx], emi = g->emission[idx];
    R albedo[3];
    for (int k = 0; k < 3; ++k) {
        uint32_t c = (alb >> (8 * k)) & 0xFFu;
        albedo[k] = sizeof(R) == 4 ? (R)srgb_to_linear(c) : (R)pow((double)c / 255.0, 2.2);
    }
    R rough = (R)((mat >> 0) & 0xFFu) / (R)255, metal = (R)((mat >> 8) & 0xFFu) / (R)255,
      occl = (R)((mat >> 16) & 0xFFu) / (R)255;
    R emis[3];
    for (int k = 0; k < 3; ++k) emis[k] = (R)((emi >> (8 * k)) & 0xFFu) / (R)255;
    Shade3<R> P = ld(g->position + idx * 3), cam = ld(p->cam_pos);
    Shade3<R> V = nrm3(sub3(cam, P));
    R F0[3];
    for (int k = 0; k < 3; ++k) F0[k] = (R)0.04 * ((R)1 - metal) + albedo[k] * metal;
    R Lo[3] = {0, 0, 0};
    R a = rough * rough, a2 = a * a;
    R rr = rough + (R)1, kk = (rr * rr) / (R)8;
    for (int l = 0; l < p->num_lights; ++l) {
        const orc_light& Lb = p->lights[l];
        Shade3<R> Lr = sub3(ld(Lb.position), P);
        Shade3<R> L = nrm3(Lr);
        R att = 1;
        Shade3<R> ndir = Shade3<R>{-(R)Lb.direction[0], -(R)Lb.direction[1], -(R)Lb.direction[2]};
        if (Lb.type == 2) {
            L = nrm3(ndir);
        } else if (Lb.type == 1) {
            R dist = len3(Lr);
            att = (R)1 / (dist * dist);
            R theta = dot3(L, nrm3(ndir));
            R eps = (R)Lb.inner_angle - (R)Lb.outer_angle;
            att = att * clamp01((theta - (R)Lb.outer_angle) / eps);
        } else {
            R dist = len3(Lr);
            att = (R)1 / (dist * dist);
        }
        R sh = shadow_factor<R>(p, shadow, l, npix, idx);
        R scale = (R)Lb.intensity * att * ((R)1 - sh);
        Shade3<R> H = nrm3(add3(V, L));
        R NdotH = mx(dot3(N, H), (R)0);
        R den = NdotH * NdotH * (a2 - (R)1) + (R)1;
        R D = a2 / (PI * den * den);
        R NdotV = mx(dot3(N, V), (R)0), NdotL = mx(dot3(N, L), (R)0);
        R G = (NdotL / (NdotL * ((R)1 - kk) + kk)) * (NdotV / (NdotV * ((R)1 - kk) + kk));
        R HV = clamp01(dot3(H, V));
        R om = clamp01((R)1 - HV);
        R om5 = om * om * om * om * om;
        R denom = (R)4 * NdotV * NdotL + (R)0.0001;
        for (int k = 0; k < 3; ++k) {
            R F = F0[k] + ((R)1 - F0[k]) * om5;
            R spec = D * G * F / denom;
            R kD = ((R)1 - F) * ((R)1 - metal);
            R radiance = (R)Lb.color[k] * scale;
            Lo[k] = Lo[k] + (kD * albedo[k] / PI + spec) * radiance * NdotL;
        }
    }
    int na = p->ao_num_samples;
    R aof = na > 0 ? (R)ao[idx] / (R)na : (R)1;
    for (int k = 0; k < 3; ++k) {
        R amb = (R)p->ambient_color[k] * (R)p->ambient_intensity * albedo[k] * occl * aof;
        out[k] = amb + Lo[k] + emis[k];
    }
    out[3] = (R)1;
}

}  // namespace

// =================================================================== C API
extern "C" {

uint32_t orc_pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

void orc_pcg4d(const uint32_t in[4], uint32_t out[4]) {
    uint32_t v[4] = {in[0], in[1], in[2], in[3]};
    pcg4d(v);
    for (int i = 0; i < 4; ++i) out[i] = v[i];
}

void orc_sample2(uint32_t x, uint32_t y, uint32_t frame, uint32_t slot, float uv[2]) {
    sample2(x, y, frame, slot, uv[0], uv[1]);
}

void orc_sincos_turns(float w, float* s, float* c) { sincos_turns(w, *s, *c); }

uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }

int orc_intersect_tri(const float o[3], const float d[3], float tmin, float tmax, const float v0[3],
                      const float v1[3], const float v2[3], float* t_out, float bary_out[3]) {
    TriRay tr = tri_setup(load3(d));
    float T, det, U, V, W;
    if (!tri_test(tr, load3(o), tmin, tmax, load3(v0), load3(v1), load3(v2), T, det, U, V, W)) return 0;
    if (t_out) *t_out = T / det;
    if (bary_out) { bary_out[0] = U / det; bary_out[1] = V / det; bary_out[2] = W / det; }
    return 1;
}

int orc_box_test(const float o[3], const float d[3], float tmin, float tmax, const float lo[3],
                 const float hi[3]) {
    BoxRay b = box_setup(load3(o), load3(d), tmin, tmax);
    float tn;
    return box_hit(b, lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], tn) ? 1 : 0;
}

void orc_invert_3x4(const float m[12], float out[12]) { invert_3x4(m, out); }

orc_scene* orc_scene_create(void) { return new orc_scene(); }
void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_scene_add_mesh(orc_scene* s, const void* vertices, uint32_t stride, uint32_t nv,
                       const uint32_t* indices, uint32_t ni) {
    if (!s || (nv && !vertices) || (ni && !indices) || ni % 3 != 0 || stride < 12) return -1;
    for (uint32_t i = 0; i < ni; ++i)
        if (indices[i] >= nv) return -1;
    Blas b;
    b.indices.assign(indices, indices + ni);
    b.vertex_count = nv;
    b.has_normals = stride >= 24;
    b.stride = stride;
    blas_build(b, vertices, stride, nv);
    s->blas.push_back(std::move(b));
    return (int)s->blas.size() - 1;
}

int orc_scene_refit_mesh(orc_scene* s, int mesh, const void* vertices) {
    if (!s || mesh < 0 || mesh >= (int)s->blas.size()) return -1;
    blas_refit(s->blas[mesh], vertices, s->blas[mesh].stride);
    tlas_build(*s);
    return 0;
}

int orc_scene_set_instances(orc_scene* s, const float* o2w, const uint32_t* mesh_ids,
                            const uint32_t* custom, const uint32_t* mask, uint32_t n) {
    if (!s) return -1;
    s->inst_input.clear();
    for (uint32_t i = 0; i < n; ++i) {
        if (mesh_ids[i] >= s->blas.size()) return -1;
        Inst in;
        std::memcpy(in.o2w, o2w + (size_t)i * 12, 48);
        invert_3x4(in.o2w, in.w2o, in.w2o_d);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) in.nmat[r * 3 + c] = in.w2o[c * 4 + r];
        in.mesh = mesh_ids[i];
        in.custom = custom ? custom[i] : i;
        in.mask = mask ? mask[i] : 0xFFu;
        in.input_index = i;
        s->inst_input.push_back(in);
    }
    tlas_build(*s);
    return 0;
}

int orc_blas_arrays(const orc_scene* s, int mesh, const void** nodes, uint32_t* node_count,
                    const float** tris, const uint32_t** prim_ids, uint32_t* tri_count) {
    if (!s || mesh < 0 || mesh >= (int)s->blas.size()) return -1;
    const Blas& b = s->blas[mesh];
    *nodes = b.nodes.data();
    *node_count = (uint32_t)b.nodes.size();
    *tris = b.tris.data();
    *prim_ids = b.prims.data();
    *tri_count = (uint32_t)b.prims.size();
    return 0;
}

int orc_tlas_arrays(const orc_scene* s, const void** nodes, uint32_t* node_count,
                    const uint32_t** inst_order, const float** w2o, uint32_t* inst_count) {
    if (!s) return -1;
    static thread_local std::vector<uint32_t> order;
    order.clear();
    for (uint32_t slot = 0; slot < s->tlas_order.size(); ++slot)
        order.push_back(s->tlas_inst[s->tlas_order[slot]].input_index);
    *nodes = s->tlas_nodes.data();
    *node_count = (uint32_t)s->tlas_nodes.size();
    *inst_order = order.data();
    *w2o = s->tlas_w2o.data();
    *inst_count = (uint32_t)order.size();
    return 0;
}

int orc_any_hit(const orc_scene* s, const float* rays, uint32_t n, uint8_t* hits, int brute_force,
                orc_stats* stats) {
    if (!s) return -1;
    orc_stats st{};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        f3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
        st.rays++;
        bool h = brute_force ? brute(*s, o, d, r[3], r[7], false, nullptr)
                             : traverse(*s, o, d, r[3], r[7], false, nullptr, &st);
        hits[i] = h ? 1 : 0;
    }
    if (stats) *stats = st;
    return 0;
}

int orc_closest_hit(const orc_scene* s, const float* rays, uint32_t n, float* t, uint32_t* inst,
                    uint32_t* prim, int brute_force) {
    if (!s) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        f3 o = mk(r[0], r[1], r[2]), d = mk(r[4], r[5], r[6]);
        Hit h{};
        bool f = brute_force ? brute(*s, o, d, r[3], r[7], true, &h)
                             : traverse(*s, o, d, r[3], r[7], true, &h, nullptr);
        t[i] = f ? h.t : INFINITY;
        inst[i] = f ? h.inst : 0xFFFFFFFFu;
        prim[i] = f ? s->blas[s->inst_input[h.inst].mesh].prims[h.slot] : 0xFFFFFFFFu;
    }
    return 0;
}

int orc_primary_gbuffer(const orc_scene* s, const orc_camera* cam, const orc_material* mats,
                        uint32_t num_mats, uint32_t W, uint32_t H, float* position, float* normal,
                        uint32_t* albedo, uint32_t* material, uint32_t* emission, int threads) {
    if (!s || !cam || W == 0 || H == 0) return -1;
    f3 eye = load3(cam->eye), fw = load3(cam->forward), rt = load3(cam->right), up = load3(cam->up);
    parallel_rows(H, threads, [&](uint32_t y, int) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t idx = (size_t)y * W + x;
            float nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
            float ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
            f3 d = mk(fw.x + nx * rt.x + ny * up.x, fw.y + nx * rt.y + ny * up.y, fw.z + nx * rt.z + ny * up.z);
            Hit h{};
            bool f = traverse(*s, eye, d, cam->tmin, cam->tmax, true, &h, nullptr);
            if (!f) {
                for (int k = 0; k < 3; ++k) { position[idx * 3 + k] = 0.0f; normal[idx * 3 + k] = 0.0f; }
                albedo[idx] = 0; material[idx] = 0; emission[idx] = 0;
                continue;
            }
            const Inst& in = s->inst_input[h.inst];
            const Blas& b = s->blas[in.mesh];
            f3 P = add(eye, mul(d, h.t));
            float bA = h.U / h.det, bB = h.V / h.det, bC = h.W / h.det;
            const float* nn = &b.nrms[(size_t)h.slot * 9];
            f3 no = mk(nn[0] * bA + nn[3] * bB + nn[6] * bC,
                       nn[1] * bA + nn[4] * bB + nn[7] * bC,
                       nn[2] * bA + nn[5] * bB + nn[8] * bC);
            const float* m = in.nmat;
            f3 nw = mk(m[0] * no.x + m[1] * no.y + m[2] * no.z,
                       m[3] * no.x + m[4] * no.y + m[5] * no.z,
                       m[6] * no.x + m[7] * no.y + m[8] * no.z);
            float ln = length(nw);
            f3 N = ln == 0.0f ? mk(0.0f, 1.0f, 0.0f) : mk(nw.x / ln, nw.y / ln, nw.z / ln);
            position[idx * 3 + 0] = P.x; position[idx * 3 + 1] = P.y; position[idx * 3 + 2] = P.z;
            normal[idx * 3 + 0] = N.x; normal[idx * 3 + 1] = N.y; normal[idx * 3 + 2] = N.z;
            if (in.custom < num_mats) {
                albedo[idx] = mats[in.custom].albedo;
                material[idx] = mats[in.custom].material;
                emission[idx] = mats[in.custom].emission;
            } else {
                albedo[idx] = 0xFFFFFFFFu;
                material[idx] = 0xFFFF0080u;  // roughness 0.5, metallic 0, occlusion 1
                emission[idx] = 0xFF000000u;
            }
        }
    });
    return 0;
}

static int trace_visibility_impl(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                                 uint8_t* shadow, uint8_t* ao, orc_stats* stats, int threads,
                                 uint32_t row_step, bool wide) {
    if (!s || !g || !p || p->num_lights < 0 || p->num_lights > 64) return -1;
    const uint32_t W = g->width, H = g->height;
    const size_t npix = (size_t)W * H;
    int nthreads = threads <= 0 ? (int)std::thread::hardware_concurrency() : threads;
    std::vector<orc_stats> per(nthreads + 1);
    parallel_rows(H, nthreads, [&](uint32_t y, int tid) {
        orc_stats& st = per[tid];  // rows not visited keep whatever the caller stored
        for (uint32_t x = 0; x < W; ++x) {
            size_t idx = (size_t)y * W + x;
            PixelRays pr = pixel_setup(g, p, idx);
            for (int l = 0; l < p->num_lights; ++l) shadow[(size_t)l * npix + idx] = 0;
            ao[idx] = 0;
            if (pr.background) continue;
            f3 P = load3(g->position + idx * 3);
            for (int l = 0; l < p->num_lights; ++l) {
                const orc_light& L = p->lights[l];
                if (traced_samples(p, l) <= 0) continue;
                ShadowSetup ss = shadow_setup(L, P);
                int cnt = 0;
                for (int smp = 0; smp < ss.n; ++smp) {
                    float u, v;
                    sample2(x, y, (uint32_t)p->frame, (uint32_t)(l * 256 + smp), u, v);
                    f3 d = shadow_dir(ss, u, v);
                    st.rays++;
                    bool h = wide ? traverse_wide(*s, pr.so, d, 0.001f, ss.tmax, &st)
                                  : traverse(*s, pr.so, d, 0.001f, ss.tmax, false, nullptr, &st);
                    if (h) cnt++;
                }
                shadow[(size_t)l * npix + idx] = (uint8_t)cnt;
            }
            int na = p->ao_num_samples;
            int miss = 0;
            for (int smp = 0; smp < na; ++smp) {
                float u, v;
                sample2(x, y, (uint32_t)p->frame, (uint32_t)(16384 + smp), u, v);
                f3 d = ao_dir(pr.aoT, pr.aoB, pr.N, u, v);
                st.rays++;
                bool h = wide ? traverse_wide(*s, pr.ao_o, d, p->ao_min, p->ao_max, &st)
                              : traverse(*s, pr.ao_o, d, p->ao_min, p->ao_max, false, nullptr, &st);
                if (!h) miss++;
            }
            ao[idx] = (uint8_t)miss;
        }
    }, row_step);
    if (stats) {
        orc_stats t{};
        for (auto& q : per) {
            t.rays += q.rays; t.node_visits += q.node_visits; t.tri_tests += q.tri_tests;
            t.inst_visits += q.inst_visits; t.hits += q.hits; t.wide_node_visits += q.wide_node_visits;
        }
        *stats = t;
    }
    return 0;
}

int orc_trace_visibility(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                         uint8_t* shadow, uint8_t* ao, orc_stats* stats, int threads,
                         uint32_t row_step) {
    return trace_visibility_impl(s, g, p, shadow, ao, stats, threads, row_step, false);
}

int orc_trace_visibility_wide(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                              uint8_t* shadow, uint8_t* ao, orc_stats* stats, int threads,
                              uint32_t row_step) {
    for (const Blas& b : s->blas)
        if (b.wide.empty()) return -2;
    return trace_visibility_impl(s, g, p, shadow, ao, stats, threads, row_step, true);
}

int orc_scene_build_wide(orc_scene* s) {
    if (!s) return -1;
    for (Blas& b : s->blas) build_wide(b.nodes, b.tris, b.prims, b.wide, b.tris8, b.prims8);
    return 0;
}

int orc_blas_wide_arrays(const orc_scene* s, int mesh, const void** nodes, uint32_t* node_count,
                         const float** tris, const uint32_t** prim_ids, uint32_t* tri_count) {
    if (!s || mesh < 0 || mesh >= (int)s->blas.size()) return -1;
    const Blas& b = s->blas[mesh];
    *nodes = b.wide.data();
    *node_count = (uint32_t)b.wide.size();
    *tris = b.tris8.data();
    *prim_ids = b.prims8.data();
    *tri_count = (uint32_t)b.prims8.size();
    return 0;
}

int orc_any_hit_wide(const orc_scene* s, const float* rays, uint32_t n, uint8_t* hits, orc_stats* stats) {
    if (!s) return -1;
    orc_stats st{};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + (size_t)i * 8;
        st.rays++;
        hits[i] = traverse_wide(*s, mk(r[0], r[1], r[2]), mk(r[4], r[5], r[6]), r[3], r[7], &st) ? 1 : 0;
    }
    if (stats) *stats = st;
    return 0;
}

// DepthToWorld (utils.glsl:1-7) for every pixel of a D32F depth image (rows
// `pitch` bytes apart): clip = (uv * 2 - 1, depth, 1) at the pixel centre
// uv = ((x + 0.5) / W, (y + 0.5) / H); v = inverse_proj * clip with the
// column products summed in column order; v /= v.w; P = (inverse_view * v).xyz.
// FP32, every operation rounded as written.
int orc_depth_to_world(const float* depth, uint32_t W, uint32_t H, uint32_t pitch, const float inv_proj[16],
                       const float inv_view[16], float* position) {
    if (!depth || !position || pitch < W * 4) return -1;
    auto mul = [](const float* m, const float v[4], float out[4]) {
        for (int r = 0; r < 4; ++r) {
            float a = m[r] * v[0];
            a = a + m[4 + r] * v[1];
            a = a + m[8 + r] * v[2];
            a = a + m[12 + r] * v[3];
            out[r] = a;
        }
    };
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            float d;
            std::memcpy(&d, reinterpret_cast<const uint8_t*>(depth) + (size_t)y * pitch + (size_t)x * 4, 4);
            const float u = ((float)x + 0.5f) / (float)W, v = ((float)y + 0.5f) / (float)H;
            const float clip[4] = {u * 2.0f - 1.0f, v * 2.0f - 1.0f, d, 1.0f};
            float vs[4], ws[4];
            mul(inv_proj, clip, vs);
            const float w = vs[3];
            for (int k = 0; k < 4; ++k) vs[k] = vs[k] / w;
            mul(inv_view, vs, ws);
            float* P = position + ((size_t)y * W + x) * 3;
            P[0] = ws[0]; P[1] = ws[1]; P[2] = ws[2];
        }
    return 0;
}

// FP64 any-hit of caller-given rays (8 doubles each: origin, tmin, direction,
// tmax) through traverse_d: the visibility decision the FP64 frame uses.
int orc_any_hit_f64(const orc_scene* s, const double* rays, uint32_t n, uint8_t* hits) {
    if (!s) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        const double* r = rays + (size_t)i * 8;
        hits[i] = traverse_d(*s, dk(r[0], r[1], r[2]), dk(r[4], r[5], r[6]), r[3], r[7]) ? 1 : 0;
    }
    return 0;
}

int orc_trace_visibility_f64(const orc_scene* s, const orc_gbuffer* g, const orc_params* p,
                             uint8_t* shadow, uint8_t* ao, int threads, uint32_t row_step) {
    if (!s || !g || !p || p->num_lights < 0 || p->num_lights > 64) return -1;
    const uint32_t W = g->width, H = g->height;
    const size_t npix = (size_t)W * H;
    const double TWO_PI = 6.283185307179586;
    parallel_rows(H, threads, [&](uint32_t y, int) {
        for (uint32_t x = 0; x < W; ++x) {
            size_t idx = (size_t)y * W + x;
            for (int l = 0; l < p->num_lights; ++l) shadow[(size_t)l * npix + idx] = 0;
            ao[idx] = 0;
            const float* pp = g->position + idx * 3; const float* pn = g->normal + idx * 3;
            d3 P = dk(pp[0], pp[1], pp[2]), N = dk(pn[0], pn[1], pn[2]);
            if (dlength(N) == 0.0) continue;
            d3 cam = dk(p->cam_pos[0], p->cam_pos[1], p->cam_pos[2]);
            double lc = dlength(dsub(P, cam));
            double sb = lc * 0.01 > 0.05 ? lc * 0.01 : 0.05;
            d3 so = dadd(P, dmul(N, sb));
            d3 aoo = dadd(P, dmul(N, lc * 0.01));
            for (int l = 0; l < p->num_lights; ++l) {
                const orc_light& L = p->lights[l];
                if (traced_samples(p, l) <= 0) continue;
                d3 lpos = dk(L.position[0], L.position[1], L.position[2]);
                d3 Lraw = dsub(lpos, P);
                double dist = dlength(Lraw);
                d3 Lvec;
                if (L.type == 2) {
                    d3 dir = dk(L.direction[0], L.direction[1], L.direction[2]);
                    d3 Ld = dnormalize(dk(-dir.x, -dir.y, -dir.z));
                    Lvec = dmul(dmul(dir, ddot(dir, Ld)), dist);
                } else {
                    Lvec = dmul(dnormalize(Lraw), dist);
                }
                d3 c = dcross(Lvec, dk(0, 1, 0));
                if (dlength(c) == 0.0) c = dcross(Lvec, dk(1, 0, 0));
                d3 T = dnormalize(c), B = dnormalize(dcross(T, Lvec));
                double tmax = dlength(Lvec);
                int cnt = 0;
                for (int smp = 0; smp < L.num_shadow_samples; ++smp) {
                    float u, v;
                    sample2(x, y, (uint32_t)p->frame, (uint32_t)(l * 256 + smp), u, v);
                    double pr = (double)L.radius * std::sqrt((double)u);
                    double th = TWO_PI * (double)v;
                    double dx = pr * std::cos(th), dy = pr * std::sin(th);
                    d3 d = dnormalize(dadd(dadd(Lvec, dmul(T, dx)), dmul(B, dy)));
                    if (traverse_d(*s, so, d, 0.001, tmax)) cnt++;
                }
                shadow[(size_t)l * npix + idx] = (uint8_t)cnt;
            }
            d3 T = std::fabs(N.z) > 0.5 ? dk(0.0, -N.z, N.y) : dk(-N.y, N.x, 0.0);
            d3 B = dcross(N, T);
            int miss = 0;
            for (int smp = 0; smp < p->ao_num_samples; ++smp) {
                float u, v;
                sample2(x, y, (uint32_t)p->frame, (uint32_t)(16384 + smp), u, v);
                double r = std::sqrt((double)u), th = 6.283 * (double)v;
                double hx = r * std::cos(th), hy = r * std::sin(th);
                double om = 1.0 - (double)u;
                double hz = std::sqrt(om > 0 ? om : 0);
                d3 d = dadd(dadd(dmul(T, hx), dmul(B, hy)), dmul(N, hz));
                if (!traverse_d(*s, aoo, d, (double)p->ao_min, (double)p->ao_max)) miss++;
            }
            ao[idx] = (uint8_t)miss;
        }
    }, row_step);
    return 0;
}

int64_t orc_visibility_rays(const orc_gbuffer* g, const orc_params* p, float* rays, int64_t cap) {
    const uint32_t W = g->width, H = g->height;
    int64_t n = 0;
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            size_t idx = (size_t)y * W + x;
            PixelRays pr = pixel_setup(g, p, idx);
            if (pr.background) continue;
            f3 P = load3(g->position + idx * 3);
            auto emit = [&](f3 o, f3 d, float tmin, float tmax) {
                if (rays && n < cap) {
                    float* r = rays + n * 8;
                    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = tmin;
                    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = tmax;
                }
                n++;
            };
            for (int l = 0; l < p->num_lights; ++l) {
                if (traced_samples(p, l) <= 0) continue;
                ShadowSetup ss = shadow_setup(p->lights[l], P);
                for (int smp = 0; smp < ss.n; ++smp) {
                    float u, v;
                    sample2(x, y, (uint32_t)p->frame, (uint32_t)(l * 256 + smp), u, v);
                    emit(pr.so, shadow_dir(ss, u, v), 0.001f, ss.tmax);
                }
            }
            for (int smp = 0; smp < p->ao_num_samples; ++smp) {
                float u, v;
                sample2(x, y, (uint32_t)p->frame, (uint32_t)(16384 + smp), u, v);
                emit(pr.ao_o, ao_dir(pr.aoT, pr.aoB, pr.N, u, v), p->ao_min, p->ao_max);
            }
        }
    return n;
}

int orc_shade(const orc_gbuffer* g, const orc_params* p, const uint8_t* shadow, const uint8_t* ao,
              float* out, int threads) {
    if (!g || !p) return -1;
    const size_t npix = (size_t)g->width * g->height;
    parallel_rows(g->height, threads, [&](uint32_t y, int) {
        for (uint32_t x = 0; x < g->width; ++x) {
            size_t idx = (size_t)y * g->width + x;
            shade_pixel<float>(g, p, shadow, ao, idx, npix, out + idx * 4);
        }
    });
    return 0;
}

int orc_shade_f64(const orc_gbuffer* g, const orc_params* p, const uint8_t* shadow,
                  const uint8_t* ao, double* out, int threads) {
    if (!g || !p) return -1;
    const size_t npix = (size_t)g->width * g->height;
    parallel_rows(g->height, threads, [&](uint32_t y, int) {
        for (uint32_t x = 0; x < g->width; ++x) {
            size_t idx = (size_t)y * g->width + x;
            shade_pixel<double>(g, p, shadow, ao, idx, npix, out + idx * 4);
        }
    });
    return 0;
}

}  // extern "C"
