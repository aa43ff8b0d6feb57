// cull.hpp — ray-group culling of visibility launches (gfx950), shared by the
// single-scene calls (trace.hip) and batched frames (batch.hip).
//
// The rays of a group of 64 consecutive active pixels (a ray group of the
// grouped order) that share a kind (the samples of one light, or the AO
// samples) lie in a region bounded cheaply from the group's pixel records:
//   shadow rays of light l run from origins o in the box A of the group's
//   shadow origins to end points within 2 x radius of o + Lvec (light.frag
//   :111-146): Lvec = the light vector, so o + Lvec is the light position
//   lifted by the origin's offset N x sb for point and spot lights, and is
//   computed per pixel for directional lights (box C of the end points).
//   Every point of such a ray is within max(|A|/2, |C|/2 + 2 radius) of the
//   segment between the centres of A and C, at the same fraction of its
//   length, so a TLAS slot whose padded world box, widened by that distance,
//   misses the segment cannot hold a hit of any ray of the group and light;
//   AO rays run from origins within 0.05 |N| of A for at most
//   ao_max x max(|N|, |N|^2) (light.frag:86-97,229-231), so a slot box that
//   misses A widened by that reach cannot hold one of their hits.
// k_cull writes, per group, the candidate mask of each kind (bit s: TLAS slot
// s) and the group's live rows (slots j whose kind has a candidate); rays of
// other rows are misses without being traced.  The widening carries a margin
// of 1e-3 + 2e-5 x the magnitude of the coordinates involved, far beyond the
// FP32 rounding that separates these bounds from the traced rays.  A group
// whose records hold a non-finite value gets every slot as candidate.
#pragma once
#include "lightpath.hpp"

namespace luz {

constexpr uint32_t kCullGroup = 64;     // pixels per culling group (= kRayGroup)
constexpr uint32_t kMaxCullInst = 128;  // TLAS slots a candidate mask covers (kCandWords x 64)
constexpr uint32_t kCandWords = 2;      // u64 words of a candidate mask
constexpr uint32_t kMaxCullSlots = 64;  // ray slots per pixel a live-row mask covers

// Wave-wide min / max of finite floats (the result, uniform, in every lane):
// each float mapped to an integer of the same order (sign and magnitude to
// two's complement), v_min_i32 / v_max_i32 with DPP exchanges within each row
// of 16 lanes (quad xor 1, quad xor 2, half-row mirror, row mirror), the four
// row values combined on the scalar unit after readlane, and mapped back.
// Exact, as min and max do not round; a non-finite input may give any value
// (cull_group takes every slot as a candidate for such a group).
__device__ inline int order_key(float v) {
    const int b = __float_as_int(v);
    return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ inline float key_value(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF)); }
template <int CTRL>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
// The bounds lo[i] = wave min, hi[i] = wave max of v[i], i < N, with the
// exchanges of all 2N values interleaved step by step (so each DPP read finds
// its operand written several instructions earlier).
template <int N>
__device__ inline void wave_bounds(const float (&v)[N], float (&lo)[N], float (&hi)[N]) {
    int a[N], b[N];
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = b[i] = order_key(v[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) { a[i] = min(a[i], dpp_i<0xB1>(a[i])); b[i] = max(b[i], dpp_i<0xB1>(b[i])); }
#pragma unroll
    for (int i = 0; i < N; ++i) { a[i] = min(a[i], dpp_i<0x4E>(a[i])); b[i] = max(b[i], dpp_i<0x4E>(b[i])); }
#pragma unroll
    for (int i = 0; i < N; ++i) { a[i] = min(a[i], dpp_i<0x141>(a[i])); b[i] = max(b[i], dpp_i<0x141>(b[i])); }
#pragma unroll
    for (int i = 0; i < N; ++i) { a[i] = min(a[i], dpp_i<0x140>(a[i])); b[i] = max(b[i], dpp_i<0x140>(b[i])); }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int a0 = __builtin_amdgcn_readlane(a[i], 0), a1 = __builtin_amdgcn_readlane(a[i], 16),
                  a2 = __builtin_amdgcn_readlane(a[i], 32), a3 = __builtin_amdgcn_readlane(a[i], 48);
        const int b0 = __builtin_amdgcn_readlane(b[i], 0), b1 = __builtin_amdgcn_readlane(b[i], 16),
                  b2 = __builtin_amdgcn_readlane(b[i], 32), b3 = __builtin_amdgcn_readlane(b[i], 48);
        lo[i] = key_value(min(min(a0, a1), min(a2, a3)));
        hi[i] = key_value(max(max(b0, b1), max(b2, b3)));
    }
}
__device__ inline bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// Does the segment p0 + s (p1 - p0), s in [0, 1], meet the box [lo - w, hi + w]?
// inv[k] = 1 / (p1[k] - p0[k]) by the hardware reciprocal (its one-ulp error
// moves the crossing points by far less than the widening's margin); a zero
// component tests p0 against the slab instead.
__device__ inline bool segment_meets_box(const float p0[3], const float p1[3], const float inv[3],
                                         const float lo[3], const float hi[3], float w) {
    float s0 = 0.0f, s1 = 1.0f;
    bool in = true;
    for (int k = 0; k < 3; ++k) {
        const float a = lo[k] - w, b = hi[k] + w;
        if (p1[k] == p0[k]) {
            in = in && p0[k] >= a && p0[k] <= b;
        } else {
            float t0 = (a - p0[k]) * inv[k], t1 = (b - p0[k]) * inv[k];
            if (t0 > t1) { const float t = t0; t0 = t1; t1 = t; }
            s0 = fmax_(s0, t0);
            s1 = fmin_(s1, t1);
        }
    }
    return in && s0 <= s1;
}

// The culling of one group, by one wave: lane i holds the frame f of the
// group's pixel i (lanes past its last pixel repeat the first one).  lights
// and slot_base are the scene's (LDS or global), slot_box and ninst its TLAS
// slots' padded world boxes (ninst <= kMaxCullInst: lane i tests slots i and
// i + 64); wave_masks is this wave's scratch of nl + 1 masks of kCandWords
// words (LDS).  Writes the nl + 1 candidate masks of the group to cand (word w
// of kind k at cand[k * kCandWords + w]) and returns its live-row bits (bit j:
// slot j < R has a candidate).
template <typename LIGHT, typename SLOT>
__device__ inline uint64_t cull_group(const PixelFrame& f, const LIGHT* lights, const SLOT* slot_base,
                                      uint32_t nl, uint32_t r_shadow, uint32_t R, float ao_max,
                                      const float* slot_box, uint32_t ninst, uint64_t* wave_masks,
                                      uint64_t* cand) {
    const uint32_t lane = threadIdx.x & 63u;
    const V3 o = shadow_origin(f);
    const float nlen = __builtin_amdgcn_sqrtf(vdot(f.N, f.N));
    // the shadow origin's offset from the surface, N x sb: a ray towards a
    // point or spot light ends near the light position plus this offset
    const V3 lv = vmul(f.N, fmax_(f.lc * 0.01f, 0.05f));
    const bool bad = !(finite3(o) && finite3(f.P) && finite3(lv) && isfinite(nlen));
    const bool any_bad = __ballot(bad) != 0ull;
    float alo[3], ahi[3], llo[3], lhi[3], nlo[1], nhi[1];
    {
        const float ov[3] = {o.x, o.y, o.z}, lvv[3] = {lv.x, lv.y, lv.z}, nv[1] = {nlen};
        wave_bounds(ov, alo, ahi);
        wave_bounds(lvv, llo, lhi);
        wave_bounds(nv, nlo, nhi);
    }
    const float nmax = nhi[0];
    float l0[3], rl2 = 0.0f;
    for (int k = 0; k < 3; ++k) {
        l0[k] = 0.5f * (llo[k] + lhi[k]);
        const float h = 0.5f * (lhi[k] - llo[k]);
        rl2 += h * h;
    }
    const float rL = __builtin_amdgcn_sqrtf(rl2) * 1.0001f;
    float a0[3], ra2 = 0.0f, mag = 0.0f;
    for (int k = 0; k < 3; ++k) {
        a0[k] = 0.5f * (alo[k] + ahi[k]);
        const float h = 0.5f * (ahi[k] - alo[k]);
        ra2 += h * h;
        mag = fmax_(mag, fmax_(fabsf(alo[k]), fabsf(ahi[k])));
    }
    const float rA = __builtin_amdgcn_sqrtf(ra2) * 1.0001f;
    // this lane's TLAS slot boxes (slots lane and lane + 64)
    const bool own = lane < ninst, own1 = lane + 64 < ninst, two = ninst > 64;
    float blo[3] = {0.0f, 0.0f, 0.0f}, bhi[3] = {0.0f, 0.0f, 0.0f};
    float blo1[3] = {0.0f, 0.0f, 0.0f}, bhi1[3] = {0.0f, 0.0f, 0.0f};
    if (own) {
        for (int k = 0; k < 3; ++k) {
            blo[k] = slot_box[lane * 6 + k];
            bhi[k] = slot_box[lane * 6 + 3 + k];
        }
    }
    if (own1) {
        for (int k = 0; k < 3; ++k) {
            blo1[k] = slot_box[(lane + 64) * 6 + k];
            bhi1[k] = slot_box[(lane + 64) * 6 + 3 + k];
        }
    }
    const uint64_t all = ninst >= 64 ? ~0ull : ((1ull << ninst) - 1ull);
    const uint64_t all1 = ninst >= 128 ? ~0ull : (ninst > 64 ? ((1ull << (ninst - 64)) - 1ull) : 0ull);
    for (uint32_t l = 0; l < nl; ++l) {
        const LIGHT& Lt = lights[l];
        uint64_t m = 0, m1 = 0;
        if (Lt.num_shadow_samples > 0) {
            const V3 lpos = v3(Lt.position[0], Lt.position[1], Lt.position[2]);
            const bool lbad = !(finite3(lpos) && isfinite(Lt.direction[0]) && isfinite(Lt.direction[1]) &&
                                isfinite(Lt.direction[2]) && isfinite(Lt.radius));
            float c0[3], rC, cmag;
            if (Lt.type == LUZRT_LIGHT_DIRECTIONAL) {
                // end points o + dir (dir . normalize(-dir)) |lpos - P| (shadow_setup)
                const V3 dir = v3(Lt.direction[0], Lt.direction[1], Lt.direction[2]);
                const float kd = vdot(dir, vnorm(v3(-dir.x, -dir.y, -dir.z)));
                const V3 e = vadd(o, vmul(dir, kd * vlen(vsub(lpos, f.P))));
                const float ev[3] = {e.x, e.y, e.z};
                float elo[3], ehi[3];
                wave_bounds(ev, elo, ehi);
                float rc2 = 0.0f;
                cmag = 0.0f;
                for (int k = 0; k < 3; ++k) {
                    c0[k] = 0.5f * (elo[k] + ehi[k]);
                    const float h = 0.5f * (ehi[k] - elo[k]);
                    rc2 += h * h;
                    cmag = fmax_(cmag, fmax_(fabsf(elo[k]), fabsf(ehi[k])));
                }
                rC = sqrtf(rc2) * 1.0001f;
            } else {
                // end points o + Lvec = P + N sb + (lpos - P): the light position
                // plus the group's offsets
                c0[0] = lpos.x + l0[0]; c0[1] = lpos.y + l0[1]; c0[2] = lpos.z + l0[2];
                rC = rL;
                cmag = fmax_(fabsf(c0[0]), fmax_(fabsf(c0[1]), fabsf(c0[2]))) + rL;
            }
            rC += 2.0f * fabsf(Lt.radius);
            const float w = fmax_(rA, rC) + 1e-3f + 2e-5f * (mag + cmag + rA + rC);
            const float inv[3] = {__builtin_amdgcn_rcpf(c0[0] - a0[0]), __builtin_amdgcn_rcpf(c0[1] - a0[1]),
                                  __builtin_amdgcn_rcpf(c0[2] - a0[2])};
            const bool meet = own && segment_meets_box(a0, c0, inv, blo, bhi, w);
            m = __ballot(meet);
            if (two) m1 = __ballot(own1 && segment_meets_box(a0, c0, inv, blo1, bhi1, w));
            if (any_bad || lbad || !isfinite(w)) {
                m = all;
                m1 = all1;
            }
        }
        if (lane == 0) {
            wave_masks[l * kCandWords] = m;
            wave_masks[l * kCandWords + 1] = m1;
        }
    }
    {
        const float amax = fmax_(ao_max, 0.0f);
        const float reach = amax * fmax_(nmax, nmax * nmax) * 1.00001f + 0.05f * nmax;
        const float w = reach + 1e-3f + 2e-5f * (mag + reach);
        bool meet = own, meet1 = own1;
        for (int k = 0; k < 3; ++k) {
            meet = meet && alo[k] - w <= bhi[k] && ahi[k] + w >= blo[k];
            meet1 = meet1 && alo[k] - w <= bhi1[k] && ahi[k] + w >= blo1[k];
        }
        uint64_t m = __ballot(meet), m1 = __ballot(meet1);
        if (any_bad || !isfinite(ao_max) || !isfinite(w)) {
            m = all;
            m1 = all1;
        }
        if (lane == 0) {
            wave_masks[nl * kCandWords] = m;
            wave_masks[nl * kCandWords + 1] = m1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = lane; k < (nl + 1) * kCandWords; k += 64) cand[k] = wave_masks[k];
    const uint32_t j = lane;
    bool live = false;
    if (j < R) {
        uint32_t kind = nl;
        if (j < r_shadow) {   // the last light whose slot base is <= j (slot_light)
            int lo = 0, hi = (int)nl;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (slot_base[mid] <= j) lo = mid; else hi = mid;
            }
            kind = (uint32_t)lo;
        }
        live = (wave_masks[kind * kCandWords] | wave_masks[kind * kCandWords + 1]) != 0ull;
    }
    return __ballot(live);
}

// Lane j of a group's wave: live row j goes to the row list at its rank among
// the group's live rows (rows[0] = the group's first list entry, code = the
// group's row code g << 6), a culled row's w hit bytes (grouped order: slot j
// of the group's w pixels, h = the group's first hit byte) are written as misses.
__device__ inline void cull_rows_lane(uint64_t bits, uint32_t R, uint32_t w, uint32_t code, uint32_t* rows,
                                      uint8_t* h) {
    const uint32_t j = threadIdx.x & 63u;
    if (j >= R) return;
    if ((bits >> j) & 1ull) {
        const uint64_t below = j ? (bits & (~0ull >> (64 - j))) : 0ull;
        rows[(uint32_t)__popcll(below)] = code | j;
    } else {
        uint8_t* hr = h + (size_t)j * w;
        if (w == kCullGroup && ((uintptr_t)hr & 15u) == 0) {
            uint4* h4 = reinterpret_cast<uint4*>(hr);
            for (int k = 0; k < 4; ++k) h4[k] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (uint32_t k = 0; k < w; ++k) hr[k] = 0;
        }
    }
}

// Exclusive scan in place of n counts (16-byte aligned) by one workgroup of
// 1024 threads: thread t sums its run of `per` consecutive counts (a multiple
// of 4, read as 16-byte vectors so that a run's loads are in flight
// together), the runs' sums are scanned by wave shuffles and one exchange of
// wave totals, and each run is rewritten with its prefixes.  Returns the total
// (to every thread).
__device__ inline uint32_t scan_runs_1024(uint32_t* counts, uint32_t n) {
    __shared__ uint32_t wave_sum[1024 / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t per = (((n + 1023u) / 1024u) + 3u) & ~3u, begin = threadIdx.x * per;
    // the run [begin, end): nv whole vectors, then single counts
    const uint32_t end = begin + per < n ? begin + per : (begin < n ? n : begin);
    const uint32_t nv = (end - begin) >> 2, tail = begin + nv * 4;
    uint4* c4 = reinterpret_cast<uint4*>(counts) + (begin >> 2);
    uint32_t sum = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < nv; ++k) {
        const uint4 v = c4[k];
        sum += v.x + v.y + v.z + v.w;
    }
    for (uint32_t q = tail; q < end; ++q) sum += counts[q];
    uint32_t incl = sum;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_sum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t w = 0; w < 1024 / 64; ++w) {
        if (w < wave) before += wave_sum[w];
        total += wave_sum[w];
    }
    uint32_t run = before + incl - sum;
#pragma unroll 8
    for (uint32_t k = 0; k < nv; ++k) {
        const uint4 v = c4[k];
        uint4 o;
        o.x = run; run += v.x;
        o.y = run; run += v.y;
        o.z = run; run += v.z;
        o.w = run; run += v.w;
        c4[k] = o;
    }
    for (uint32_t q = tail; q < end; ++q) {
        const uint32_t v = counts[q];
        counts[q] = run;
        run += v;
    }
    return total;
}

}  // namespace luz
