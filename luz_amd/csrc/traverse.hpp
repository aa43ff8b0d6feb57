// traverse.hpp — two-level BVH traversal state machine for gfx950.
//
// One lane owns one ray.  Its stack lives in LDS (kShort entries per lane,
// interleaved so that lane i of a wave touches bank i) and spills its oldest
// entries to a per-lane global array when it overflows, so the depth is
// unbounded by LDS size.  A sentinel entry on the stack marks the return from
// a BLAS to the TLAS (the instance's object-space ray is then dropped).
//
// The step function advances a ray by one node visit (two slab tests, and the
// triangle tests of leaf children) so that a persistent kernel can interleave
// steps with wave-wide refills of finished lanes.
#pragma once
#include "device_math.hpp"

namespace luz {

constexpr int kShort = 16;    // default LDS stack entries per lane (power of two)
constexpr int kSpill = 160;   // global overflow entries per lane (>= TLAS+BLAS depth + 1)

typedef uint32_t __attribute__((address_space(3))) LdsWord;
typedef uint32_t __attribute__((address_space(1))) GlobalWord;

// SHORT: LDS entries per lane (a power of two, chosen by the autotuner);
// DEPTH: the deepest stack a ray may use (its spill column holds DEPTH
// entries).  The two columns carry their address spaces, so a pop reads LDS
// with an LDS load and only an overflowed entry with a global load (a generic
// pointer would turn every pop into a flat load that waits on both memory
// counters).
template <int SHORT, int DEPTH = kSpill>
struct StackT {
    LdsWord* ls;      // this lane's LDS column; entry e at ls[e * lds_stride]
    GlobalWord* gs;   // this lane's spill column; entry e at gs[e * g_stride]
    uint32_t lds_stride;
    uint32_t g_stride;
    int sp, lo;

    __device__ inline void init(uint32_t* lds_column, uint32_t lds_step, uint32_t* spill_column,
                                uint32_t spill_step) {
        ls = (LdsWord*)lds_column;
        gs = (GlobalWord*)spill_column;
        lds_stride = lds_step;
        g_stride = spill_step;
        sp = 0;
        lo = 0;
    }

    // A full stack (DEPTH entries) refuses the push before anything spills, so
    // the spill column is written at entries 0 .. DEPTH - SHORT - 1 only.
    __device__ inline bool push(uint32_t x) {
        if (sp >= DEPTH) return false;
        if (sp - lo == SHORT) {
            gs[(size_t)lo * g_stride] = ls[(lo & (SHORT - 1)) * lds_stride];
            lo++;
        }
        ls[(sp & (SHORT - 1)) * lds_stride] = x;
        sp++;
        return true;
    }
    __device__ inline uint32_t pop() {
        sp--;
        if (sp < lo) {
            lo = sp;
            return gs[(size_t)sp * g_stride];
        }
        return ls[(sp & (SHORT - 1)) * lds_stride];
    }
};
using Stack = StackT<kShort>;

struct Best {
    float t;
    uint32_t inst_slot, tri_slot;
    float T, det, U, V, W;
    bool found;
};

struct RayState {
    V3 wd;         // world direction (the world origin is wbr.o)
    float tmin, tmax;
    BoxRay br;     // the object-space slab-test ray of the instance being walked
    BoxRay wbr;    // the world-space one (TLAS nodes)
    TriRay tr;
    const Node* bnodes;
    const float* btris;
    uint32_t inst_slot;
    uint32_t next;
    uint32_t steps;
    bool in_blas;
};

// A well-formed two-level tree visits each node at most once per instance
// entry; the cap only guards against a malformed structure looping forever.
constexpr uint32_t kMaxSteps = 1u << 24;

enum StepResult { kStepContinue = 0, kStepHit = 1, kStepDone = 2, kStepOverflow = 3 };

template <typename STACK>
__device__ inline void ray_begin(RayState& r, STACK& st, V3 o, V3 d, float tmin, float tmax) {
    r.wd = d; r.tmin = tmin; r.tmax = tmax;
    r.wbr = box_setup(o, d, tmin, tmax);
    r.br = r.wbr;
    r.next = 0;
    r.steps = 0;
    r.in_blas = false;
    st.sp = 0; st.lo = 0;
}

typedef float Row4 __attribute__((ext_vector_type(4)));
typedef const Row4 __attribute__((address_space(1)))* GlobalRow;
typedef const Row4 __attribute__((address_space(3)))* LdsRow;

// Four 16-byte rows of entry i (a node or an instance record, 64 bytes) of an
// array in global memory or, when `lds`, of a TLAS staged in LDS.  Naming the
// address space lets the compiler issue four 16-byte loads instead of generic
// per-dword loads; passing base and index (not a selected pointer) keeps the
// LDS base an LDS address.
template <typename T>
__device__ inline void load_rows(const T* base, uint32_t i, bool lds, float4& a, float4& b, float4& c,
                                 uint4& d) {
    static_assert(sizeof(T) == 64, "64-byte entries");
    Row4 r[4];
    if (lds) {
        LdsRow p = (LdsRow)(base + i);
        r[0] = p[0]; r[1] = p[1]; r[2] = p[2]; r[3] = p[3];
    } else {
        GlobalRow p = (GlobalRow)(base + i);
        r[0] = p[0]; r[1] = p[1]; r[2] = p[2]; r[3] = p[3];
    }
    a = make_float4(r[0].x, r[0].y, r[0].z, r[0].w);
    b = make_float4(r[1].x, r[1].y, r[1].z, r[1].w);
    c = make_float4(r[2].x, r[2].y, r[2].z, r[2].w);
    d = make_uint4(__float_as_uint(r[3].x), __float_as_uint(r[3].y), __float_as_uint(r[3].z),
                   __float_as_uint(r[3].w));
}

// Triangle slot of a BLAS (9 floats, global memory: global loads, not flat).
__device__ inline void load_tri(const float* tris, uint32_t slot, V3& a, V3& b, V3& c) {
    typedef const float __attribute__((address_space(1)))* GlobalFloat;
    const GlobalFloat t = (GlobalFloat)(tris + (size_t)slot * 9);
    a = v3(t[0], t[1], t[2]); b = v3(t[3], t[4], t[5]); c = v3(t[6], t[7], t[8]);
}

// Closest-hit ordering: smaller t, then smaller instance input index, then primitive id.
__device__ inline bool better_hit(float t, uint32_t inst_slot, uint32_t tri_slot, const Best& b,
                                  const InstAux* aux) {
    if (!b.found || t < b.t) return true;
    if (t > b.t) return false;
    uint32_t ia = aux[inst_slot].input_index, ib = aux[b.inst_slot].input_index;
    if (ia != ib) return ia < ib;
    return aux[inst_slot].prims[tri_slot] < aux[b.inst_slot].prims[b.tri_slot];
}

__device__ inline uint32_t byte_of(uint32_t lo, uint32_t hi, int k) {
    return ((k < 4 ? lo : hi) >> (8 * (k & 3))) & 0xFFu;
}

// Visit of an 8-wide quantised BLAS node (any-hit): 8 slab tests on the
// decoded child boxes, then the hit slots in slot order.
template <typename STACK>
__device__ inline int wide_visit(RayState& r, STACK& st) {
    const uint4* p = reinterpret_cast<const uint4*>(reinterpret_cast<const Node8*>(r.bnodes) + r.next);
    uint4 h0 = p[0], h1 = p[1], qa = p[2], qb = p[3], qc = p[4];
    const float ox = __uint_as_float(h0.x), oy = __uint_as_float(h0.y), oz = __uint_as_float(h0.z);
    const int ex = (int)(int8_t)(h0.w & 0xFFu), ey = (int)(int8_t)((h0.w >> 8) & 0xFFu),
              ez = (int)(int8_t)((h0.w >> 16) & 0xFFu);
    const int nchild = (int)(h0.w >> 24);
    const float sx = __int_as_float((ex + 127) << 23), sy = __int_as_float((ey + 127) << 23),
                sz = __int_as_float((ez + 127) << 23);
    uint32_t mask = 0;
    // the near and far planes' byte words per axis, chosen once per node by
    // the ray's direction signs (box_hit chooses them per child)
    const bool gx = r.br.inv.x < 0.0f, gy = r.br.inv.y < 0.0f, gz = r.br.inv.z < 0.0f;
    const uint32_t nxa = gx ? qb.z : qa.x, nxb = gx ? qb.w : qa.y, fxa = gx ? qa.x : qb.z, fxb = gx ? qa.y : qb.w;
    const uint32_t nya = gy ? qc.x : qa.z, nyb = gy ? qc.y : qa.w, fya = gy ? qa.z : qc.x, fyb = gy ? qa.w : qc.y;
    const uint32_t nza = gz ? qc.z : qb.x, nzb = gz ? qc.w : qb.y, fza = gz ? qb.x : qc.z, fzb = gz ? qb.y : qc.w;
    // A child plane o + q x 2^e is met at t = (o + q x 2^e - ro) x inv =
    // q x S + O, with S = 2^e x inv (exact: a power-of-two scaling) and
    // O = (o - ro) x inv per node and axis: one fused multiply-add per plane.
    // Its rounding error, like that of decoding the plane first, is a few
    // ulps of the terms, covered as before by the far bound's widening and
    // the child boxes' outward rounding; an infinite reciprocal (a direction
    // component of zero) gives infinite or NaN distances that the max/min
    // ignore or that keep the box, as before.
    const float Sx = sx * r.br.inv.x, Sy = sy * r.br.inv.y, Sz = sz * r.br.inv.z;
    const float Ox = (ox - r.br.o.x) * r.br.inv.x, Oy = (oy - r.br.o.y) * r.br.inv.y,
                Oz = (oz - r.br.o.z) * r.br.inv.z;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float tnx = __builtin_fmaf((float)byte_of(nxa, nxb, k), Sx, Ox);
        const float tfx = __builtin_fmaf((float)byte_of(fxa, fxb, k), Sx, Ox);
        const float tny = __builtin_fmaf((float)byte_of(nya, nyb, k), Sy, Oy);
        const float tfy = __builtin_fmaf((float)byte_of(fya, fyb, k), Sy, Oy);
        const float tnz = __builtin_fmaf((float)byte_of(nza, nzb, k), Sz, Oz);
        const float tfz = __builtin_fmaf((float)byte_of(fza, fzb, k), Sz, Oz);
        if (k < nchild && box_hit_t(r.br, tnx, tfx, tny, tfy, tnz, tfz)) mask |= 1u << k;
    }
    const uint64_t meta = ((uint64_t)h1.w << 32) | h1.z;
    r.next = kNone;
    while (mask) {
        int k = __ffs(mask) - 1;
        mask &= mask - 1;
        uint32_t m = (uint32_t)(meta >> (8 * k)) & 0xFFu;
        if (m & 0x80u) {
            uint32_t child = h1.x + (m & 0x7Fu);
            if (r.next == kNone) r.next = child;
            else if (!st.push(child)) return kStepOverflow;
        } else {
            V3 a, b, c;
            load_tri(r.btris, h1.y + m, a, b, c);
            float T, det, U, V, W;
            if (tri_test(r.tr, r.br.o, r.tmin, r.tmax, a, b, c, T, det, U, V, W)) return kStepHit;
        }
    }
    return kStepContinue;
}

// tlas_lds: tlas_nodes and recs point into LDS (a staged TLAS).
// TLAS_FREE: the ray visits no TLAS node (its stack holds instance slots and
// BLAS nodes only, as after ray-group culling), so the world slab ray is not
// kept and every node visited is a BLAS node.
template <bool CLOSEST, bool WIDE = false, typename STACK = Stack, bool TLAS_FREE = false>
__device__ inline int ray_step(RayState& r, STACK& st, const Node* tlas_nodes, const InstRec* recs,
                               const InstAux* aux, Best* best, bool tlas_lds = false,
                               uint32_t max_steps = kMaxSteps) {
    if (++r.steps > max_steps) return kStepOverflow;
    if (r.next == kNone) {
        if (st.sp == 0) return kStepDone;
        r.next = st.pop();
        if (r.next == kSentinel) {
            r.in_blas = false;
            r.next = kNone;
            return kStepContinue;
        }
    }
    // a TLAS-free ray holds only instance slots at the top level
    if (TLAS_FREE && !r.in_blas && !(r.next & kLeaf)) return kStepOverflow;
    if (!r.in_blas && (r.next & kLeaf)) {
        uint32_t slot = r.next & ~kLeaf;
        float4 w0, w1, w2;
        uint4 pp;
        load_rows(recs, slot, tlas_lds, w0, w1, w2, pp);
        float w[12] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
        V3 oo = xform_point(w, r.wbr.o);
        V3 od = xform_vec(w, r.wd);
        float thi = CLOSEST ? best->t : r.tmax;
        r.br = box_setup(oo, od, r.tmin, thi);
        r.tr = tri_setup(od);
        r.bnodes = reinterpret_cast<const Node*>(((uint64_t)pp.y << 32) | pp.x);
        r.btris = reinterpret_cast<const float*>(((uint64_t)pp.w << 32) | pp.z);
        r.inst_slot = slot;
        if (!st.push(kSentinel)) return kStepOverflow;
        r.in_blas = true;
        r.next = 0;
        return kStepContinue;
    }
    if (WIDE && (TLAS_FREE || r.in_blas)) return wide_visit(r, st);
    float4 qx, qy, qz;
    uint4 qc;
    if (TLAS_FREE)
        load_rows(r.bnodes, r.next, false, qx, qy, qz, qc);
    else
        load_rows(r.in_blas ? r.bnodes : tlas_nodes, r.next, tlas_lds && !r.in_blas, qx, qy, qz, qc);
    float tn0, tn1;
    // the ray of the tree being walked, selected per lane (restoring the world
    // ray into a single slot on every return to the TLAS costs more)
    const BoxRay cur = (TLAS_FREE || r.in_blas) ? r.br : r.wbr;
    const bool b0 = box_hit(cur, qx.x, qx.z, qy.x, qy.z, qz.x, qz.z, tn0);
    const bool b1 = box_hit(cur, qx.y, qx.w, qy.y, qy.w, qz.y, qz.w, tn1);
    const bool h0 = (qc.x != kEmpty) & b0;
    const bool h1 = (qc.y != kEmpty) & b1;
    int first = (h0 && h1) ? (tn1 < tn0 ? 1 : 0) : (h0 ? 0 : 1);
    r.next = kNone;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int c = k == 0 ? first : 1 - first;
        bool h = c == 0 ? h0 : h1;
        if (!h) continue;
        uint32_t ch = c == 0 ? qc.x : qc.y;
        if ((TLAS_FREE || r.in_blas) && (ch & kLeaf)) {
            uint32_t slot = ch & ~kLeaf;
            V3 a, b, cc;
            load_tri(r.btris, slot, a, b, cc);
            float T, det, U, V, W;
            if (tri_test(r.tr, r.br.o, r.tmin, r.tmax, a, b, cc, T, det, U, V, W)) {
                if (!CLOSEST) return kStepHit;
                float t = __fdiv_rn(T, det);
                if (better_hit(t, r.inst_slot, slot, *best, aux)) {
                    best->found = true;
                    best->t = t;
                    best->inst_slot = r.inst_slot;
                    best->tri_slot = slot;
                    best->T = T; best->det = det; best->U = U; best->V = V; best->W = W;
                    float m = t_margin(t);
                    r.br.thi = t + m;
                    r.wbr.thi = t + m;
                }
            }
            continue;
        }
        if (r.next == kNone) r.next = ch;
        else if (!st.push(ch)) return kStepOverflow;
    }
    return kStepContinue;
}

}  // namespace luz
