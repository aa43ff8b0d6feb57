// trace.hip — shadow and ambient-occlusion visibility on gfx950.
//
// Replaces the ray queries of light.frag (TraceShadowRay :86-109,
// TraceAORays :111-135, EvaluateShadow :137-146, AO origin :229-231) with a
// persistent traversal over the active pixels:
//   k_compact_* non-background pixels in 64-pixel-tile order -> active list
//               (count, scan, write: deterministic, no atomics); the scan
//               also sets the call's counters (active pixels, rays = active
//               pixels x slots, claim and error counters zeroed);
//   k_trace_any persistent any-hit traversal: a wave refills its idle lanes
//               with one atomic when at least `refill` lanes are idle
//               (ballot-based compaction of the wave's work); each ray
//               (ray_row order: a wave holds one light sample or AO sample
//               of 64 neighbouring pixels) is generated at claim time from
//               the 32-byte record of its pixel that the compaction wrote
//               (gen_ray), so no ray records go through memory (they remain
//               for luzrt_trace_rays);
//   k_resolve   per active pixel (in active order), occluded shadow samples
//               per light and unoccluded AO samples, written as u8 planes;
//               background pixels get zero counts.
#include <algorithm>

#include "internal.hpp"
#include "lightpath.hpp"
#include "traverse.hpp"
#include "cull.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

// Counters of a trace_rays call; the error latch kCntError is left as it is
// (only the host clears it, when it reports it).
__global__ void k_vis_init(uint32_t* counters, uint32_t ray_count) {
    if (threadIdx.x <= kCntRays && threadIdx.x != kCntError)
        counters[threadIdx.x] = threadIdx.x == kCntRays ? ray_count : 0u;
    clear_claims(counters, threadIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_compact_count(const luzrt_gbuffer gb, uint32_t tw,
                                                          uint32_t ntiles_px, uint32_t* block_counts) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    uint32_t p;
    bool inside;
    bool active = pixel_active(gb, blockIdx.x * kBlock + threadIdx.x, tw, ntiles_px, p, inside);
    uint32_t total;
    block_prefix(active, wave_tot, total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// Exclusive scan of the block counts.  Writes the active-pixel and ray counts
// and zeroes the claim and error counters of the traversal that follows.
__global__ __launch_bounds__(1024) void k_compact_scan(uint32_t* block_counts, uint32_t nblocks,
                                                       uint32_t* counters, const DevScene* ds) {
    const uint32_t total = scan_runs_1024(block_counts, nblocks);
    if (threadIdx.x == 0) {
        counters[kCntActive] = total;
        counters[kCntRays] = total * ds->r_total;
        counters[kCntFetch] = 0;
    }
    clear_claims(counters, threadIdx.x);
}

// Writes the active list, the per-pixel active index and the traversal's
// per-active-pixel records (pixel_record).
__global__ __launch_bounds__(kBlock) void k_compact_write(const luzrt_gbuffer gb, uint32_t tw,
                                                          uint32_t ntiles_px, const DevScene* ds,
                                                          const uint32_t* block_offsets,
                                                          uint32_t* act_list, uint32_t* act_index,
                                                          float4* act_rec) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    uint32_t p;
    bool inside;
    bool active = pixel_active(gb, g, tw, ntiles_px, p, inside);
    uint32_t total;
    uint32_t a = block_offsets[blockIdx.x] + block_prefix(active, wave_tot, total);
    if (active) {
        act_list[a] = p;
        float4 r0, r1;
        pixel_record(pixel_frame(gb, ds->s, p), r0, r1);
        act_rec[2 * (size_t)a] = r0;
        act_rec[2 * (size_t)a + 1] = r1;
    }
    if (inside) act_index[p] = active ? a : kNone;
}

// Ray-group culling (cull.hpp): one wave per 64-pixel group (four per
// workgroup) runs cull_group on the group's pixel records, with the scene's
// lights and slot bases staged in LDS once per workgroup (the per-light loop
// then reads LDS instead of waiting on scalar loads).
__global__ __launch_bounds__(kBlock) void k_cull(const float4* act_rec, const uint32_t* counters,
                                                 const DevScene* ds, const float* slot_box,
                                                 uint32_t ninst, uint32_t ngroups, uint32_t kinds,
                                                 uint64_t* cand, uint64_t* live_bits,
                                                 uint32_t* live_count) {
    __shared__ uint64_t s_mask[kBlock / 64][(LUZRT_MAX_LIGHTS + 1) * kCandWords];
    __shared__ luzrt_light s_light[LUZRT_MAX_LIGHTS];
    __shared__ uint32_t s_slot[LUZRT_MAX_LIGHTS];
    {
        const uint32_t nlw = (uint32_t)ds->s.num_lights * (uint32_t)(sizeof(luzrt_light) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(ds->s.lights);
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_light);
        for (uint32_t i = threadIdx.x; i < nlw; i += kBlock) dst[i] = src[i];
        for (uint32_t i = threadIdx.x; i < (uint32_t)ds->s.num_lights; i += kBlock) s_slot[i] = ds->slot_base[i];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (kBlock / 64) + wave;
    if (g >= ngroups) return;
    const uint32_t A = counters[kCntActive];
    if ((uint64_t)g * kCullGroup >= A) {
        if (lane == 0) {
            live_bits[g] = 0;
            live_count[g] = 0;
        }
        return;
    }
    const uint32_t a = g * kCullGroup + lane;
    // lanes past the last active pixel repeat the group's first pixel
    const uint32_t ar = a < A ? a : g * kCullGroup;
    const PixelFrame f = pixel_from_record(act_rec[2 * (size_t)ar], act_rec[2 * (size_t)ar + 1]);
    const uint64_t bits = cull_group(f, s_light, s_slot, (uint32_t)ds->s.num_lights, ds->r_shadow, ds->r_total,
                                     ds->s.ao_max, slot_box, ninst, s_mask[wave],
                                     cand + (size_t)g * kinds * kCandWords);
    if (lane == 0) {
        live_bits[g] = bits;
        live_count[g] = (uint32_t)__popcll(bits);
    }
}

// Exclusive scan of the groups' live-row counts; the total goes to
// counters[kCntLive].
__global__ __launch_bounds__(1024) void k_cull_scan(uint32_t* counts, uint32_t n, uint32_t* counters) {
    const uint32_t total = scan_runs_1024(counts, n);
    if (threadIdx.x == 0) counters[kCntLive] = total;
}

// One wave per group: its live rows into the row list, its culled rows'
// hit bytes written as misses (cull_rows_lane).
__global__ __launch_bounds__(kBlock) void k_cull_rows(const uint32_t* counters, const DevScene* ds,
                                                      uint32_t ngroups, const uint64_t* live_bits,
                                                      const uint32_t* row_base, uint32_t* rows,
                                                      uint8_t* hits) {
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (kBlock / 64) + wave;
    if (g >= ngroups) return;
    const uint32_t A = counters[kCntActive];
    if ((uint64_t)g * kCullGroup >= A) return;
    const uint32_t R = ds->r_total, g0 = g * kCullGroup;
    const uint32_t w = A - g0 < kCullGroup ? A - g0 : kCullGroup;
    cull_rows_lane(live_bits[g], R, w, g << 6, rows + row_base[g], hits + (size_t)g0 * R);
}

// Rays are claimed in chunks per wave (one device-scope atomic per chunk).
// The chunk follows the launch's ray count, 1/kChunkShare of an even share per
// wave, kept within kChunkMin .. kChunkMax rays: the rays in flight on the
// whole GPU (waves x chunk) then cover few image rows of one light at a time,
// so the BVH nodes they read stay in L2, while the claims stay below the rate
// the counters sustain.  Measured on 1080p sphere fields with a single counter
// (trace ms, 2048 / 512 / 256-ray chunks): 1 shadow ray per light + 1 AO ray
// 4.40 / 2.11 / 1.82; 1 shadow ray per light + 16 AO rays 3.13 / 2.10 / 2.55;
// chunks of 64 or 128 lost again to the counter (3.6 and 2.3 ms at 1 + 1).
constexpr uint32_t kChunkMin = 256, kChunkMax = 512, kChunkShare = 12;

__device__ inline uint32_t claim_chunk(uint32_t nrays, uint32_t nwaves) {
    const uint32_t c = (nrays / (nwaves * kChunkShare)) & ~63u;
    return c < kChunkMin ? kChunkMin : (c > kChunkMax ? kChunkMax : c);
}
constexpr uint32_t kTlasLdsEntries = 240;   // TLAS nodes + instance records staged in LDS
constexpr bool kStageTlas = true;

// Scene of ray `rid` in a multi-scene launch: the last s < n with base[s] <= rid.
__device__ inline uint32_t scene_of_ray(const uint32_t* base, uint32_t n, uint32_t rid) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= rid) lo = mid; else hi = mid;
    }
    return lo;
}

// Where a launch's rays come from.  kRays: ray records (origin, tmin;
// direction, tmax).  kGen: computed at claim time from the G-buffer and the
// scene block (gen_ray).  kBatch: likewise for many scenes in one launch,
// each ray generated from and traced against its own scene (found from the
// ray index at claim time; the per-lane TLAS pointers then live in vector
// registers).
enum RaySource { kRays = 0, kBatch = 1, kGen = 2 };

// A culled ray's start: the candidate slots (mask words cw) whose padded
// boxes it meets, the test the TLAS node above each slot would make; the
// first is entered now, the others from the stack, and no TLAS node is visited.
template <typename STACK>
__device__ inline void enter_candidates(RayState& r, STACK& st, const uint64_t* cw, const float* slot_box) {
    r.next = kNone;
    for (uint32_t w = 0; w < kCandWords; ++w) {
        uint64_t m = cw[w];
        while (m) {
            const uint32_t sl = 64u * w + (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const float* b = slot_box + sl * 6;
            float tn;
            if (box_hit(r.wbr, b[0], b[3], b[1], b[4], b[2], b[5], tn)) {
                if (r.next == kNone) r.next = kLeaf | sl;
                else st.push(kLeaf | sl);
            }
        }
    }
}

// CULLED (gen launches after k_cull): the live rows only, each ray starting
// from its group's candidate slots; no TLAS node is visited, so only the
// instance records are staged in LDS.
template <bool WIDE, int SHORT, int SRC, int DEPTH = kSpill, bool CULLED = false>
__global__ __launch_bounds__(kBlock, SHORT > 16 ? 1 : 5) void k_trace_any(TraceLaunch L,
                                                                                      uint32_t refill) {
    constexpr bool BATCH = SRC == kBatch;
    __shared__ uint32_t lds[SHORT * kBlock];
    __shared__ uint32_t sbase[BATCH ? kMaxBatch + 1 : 1];
    __shared__ uint32_t sgbase[BATCH && CULLED ? kMaxBatch + 1 : 1];   // first culling group per scene
    // A small TLAS (nodes + instance records, 64 B each) is staged in LDS: every
    // ray starts there, so its reads stop competing with BLAS data in L1.  15 KB
    // keeps five 256-lane workgroups (16-entry stacks) resident per CU.
    __shared__ float4 stlas[BATCH ? 4 : (CULLED ? kMaxCullInst * 4 : kTlasLdsEntries * 4)];
    const uint32_t tid = threadIdx.x;
    if (BATCH) {
        for (uint32_t i = tid; i <= L.nscenes; i += kBlock) {
            sbase[i] = L.scene_ray_base[i];
            if (CULLED) sgbase[i] = L.scene_group_base[i];
        }
        __syncthreads();
    }
    const Node* tnodes = L.tlas_nodes;
    const InstRec* trecs = L.recs;
    bool tlas_lds = false;
    if (kStageTlas && !BATCH &&
        (CULLED ? L.recs_count <= kMaxCullInst : L.tlas_nodes_count + L.recs_count <= kTlasLdsEntries)) {
        const float4* gn = reinterpret_cast<const float4*>(L.tlas_nodes);
        const float4* gr = reinterpret_cast<const float4*>(L.recs);
        const uint32_t nn = CULLED ? 0u : L.tlas_nodes_count * 4, nr = L.recs_count * 4;
        for (uint32_t i = tid; i < nn; i += kBlock) stlas[i] = gn[i];
        for (uint32_t i = tid; i < nr; i += kBlock) stlas[nn + i] = gr[i];
        __syncthreads();
        tnodes = reinterpret_cast<const Node*>(stlas);
        trecs = reinterpret_cast<const InstRec*>(stlas + nn);
        tlas_lds = true;
    }
    const uint32_t lane = tid & 63u;
    const uint32_t total = gridDim.x * kBlock;
    StackT<SHORT, DEPTH> st;
    st.init(lds + tid, kBlock, L.spill + (size_t)blockIdx.x * kBlock + tid, total);
    // culled launches trace the live rows only (64 ray ids per row)
    constexpr bool cull = CULLED && SRC == kGen;
    constexpr bool bcull = CULLED && BATCH;
    const uint32_t nrays = CULLED ? L.counters[kCntLive] * kCullGroup : L.counters[kCntRays];
    const uint32_t A = SRC == kGen ? L.counters[kCntActive] : 1u;   // active pixels
    const uint32_t R = SRC == kGen ? L.ds->r_total : 1u;             // ray slots per pixel
    const uint32_t G = SRC == kGen ? L.ds->ray_group : 0u;           // pixels per ray group
    const uint32_t chunk = claim_chunk(nrays, total / 64u);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    RayState r;
    uint32_t ray = kNone;
    bool open = true;
    uint32_t chunk_next = 0, chunk_end = 0;  // wave-uniform
    // Claim counter c hands out chunks c, c + 8, c + 16, ... of the launch's
    // rays, so the waves still advance one front through the rays (the BVH
    // nodes in use stay few) while their atomics spread over 8 addresses.  A
    // wave starts on the counter of its workgroup and moves to the next one
    // when a counter's chunks run past the last ray.
    const uint32_t seg0 = blockIdx.x % kClaimSegments;
    uint32_t seg = 0;   // counters this wave has found exhausted (wave-uniform)
    for (;;) {
        if (open) {
            uint64_t idle = __ballot(ray == kNone);
            uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= refill || nidle == 64u) {
                while (chunk_next >= chunk_end && open) {
                    if (seg == kClaimSegments) {
                        open = false;
                        break;
                    }
                    const uint32_t c = (seg0 + seg) % kClaimSegments;
                    uint32_t got = 0;
                    if (lane == 0) got = atomicAdd(&L.counters[kClaimBase + c * kClaimStride], 1u);
                    // only lane 0 holds the atomic's result: every decision
                    // below uses the broadcast copy, so the claim stays uniform
                    const uint64_t base = ((uint64_t)__shfl(got, 0) * kClaimSegments + c) * chunk;
                    if (base < nrays) {
                        chunk_next = (uint32_t)base;
                        chunk_end = nrays - chunk_next > chunk ? chunk_next + chunk : nrays;
                    } else {
                        ++seg;
                    }
                }
                uint32_t avail = chunk_end > chunk_next ? chunk_end - chunk_next : 0u;
                uint32_t take = nidle < avail ? nidle : avail;
                if (ray == kNone) {
                    uint32_t rank = (uint32_t)__popcll(idle & lt);
                    if (rank < take) {
                        uint32_t rid = chunk_next + rank;
                        ray = rid;
                        if (bcull) {
                            // live row g << 6 | j of the batch's group space: the scene
                            // from the group bases, the hit byte at the ray's index
                            // in the scene's grouped order
                            const uint32_t row = L.live_rows[rid >> 6], i = rid & 63u;
                            const uint32_t g = row >> 6, slot = row & 63u;
                            const uint32_t sc = scene_of_ray(sgbase, L.nscenes, g);
                            const BatchTlas& bt = L.scene_tlas[sc];
                            const uint32_t As = L.scene_active[sc], g0 = (g - sgbase[sc]) * kCullGroup;
                            const uint32_t a = g0 + i;
                            if (a >= As) {
                                ray = kNone;   // past the last pixel of a partial group
                            } else {
                                const DevScene* dsc = bt.ds;
                                const uint32_t Rs = dsc->r_total, rs = dsc->r_shadow;
                                const uint32_t w = As - g0 < kCullGroup ? As - g0 : kCullGroup;
                                ray = sbase[sc] + g0 * Rs + slot * w + i;
                                trecs = bt.recs;
                                const int l = slot < rs ? slot_light(dsc, slot) : 0;
                                const uint32_t kind = slot < rs ? (uint32_t)l : (uint32_t)dsc->s.num_lights;
                                const uint64_t* cw = L.cand + ((size_t)g * L.cand_kinds + kind) * kCandWords;
                                V3 o, d;
                                float tmin, tmax;
                                gen_ray_of(dsc, pixel_frame(bt.position, bt.normal, bt.width, dsc->s,
                                                            bt.act_list[a]),
                                           slot, l, o, d, tmin, tmax);
                                ray_begin(r, st, o, d, tmin, tmax);
                                enter_candidates(r, st, cw, bt.slot_box);
                            }
                        } else if (BATCH) {
                            const uint32_t sc = scene_of_ray(sbase, L.nscenes, rid);
                            const BatchTlas& bt = L.scene_tlas[sc];
                            tnodes = bt.nodes;
                            trecs = bt.recs;
                            if (!tnodes) {   // empty scene: nothing can be hit
                                L.hits[rid] = 0;
                                ray = kNone;
                            } else {
                                uint32_t slot, a;
                                ray_slot(rid - sbase[sc], L.scene_active[sc], bt.ds->r_total,
                                         bt.ds->ray_group, slot, a);
                                V3 o, d;
                                float tmin, tmax;
                                gen_ray(bt.ds, pixel_frame(bt.position, bt.normal, bt.width, bt.ds->s,
                                                           bt.act_list[a]),
                                        slot, o, d, tmin, tmax);
                                ray_begin(r, st, o, d, tmin, tmax);
                            }
                        } else {
                            if (SRC == kGen) {
                                uint32_t slot = 0, a = 0;
                                const uint64_t* cw = nullptr;
                                const uint32_t rs = L.ds->r_shadow;
                                if (cull) {
                                    // live row g << 6 | j: slot j of the group's pixel i; its hit
                                    // byte is the ray's grouped-order index
                                    const uint32_t row = L.live_rows[rid >> 6], i = rid & 63u;
                                    const uint32_t g = row >> 6, g0 = g * kCullGroup;
                                    slot = row & 63u;
                                    a = g0 + i;
                                    if (a < A) {
                                        const uint32_t w = A - g0 < kCullGroup ? A - g0 : kCullGroup;
                                        ray = g0 * R + slot * w + i;
                                        const uint32_t kind = slot < rs ? (uint32_t)slot_light(L.ds, slot)
                                                                        : (uint32_t)L.ds->s.num_lights;
                                        cw = L.cand + ((size_t)g * L.cand_kinds + kind) * kCandWords;
                                    } else {
                                        ray = kNone;   // past the last pixel of a partial group
                                    }
                                } else {
                                    ray_slot(rid, A, R, G, slot, a);
                                }
                                if (ray != kNone) {
                                    const float4 r0 = L.act_rec[2 * (size_t)a], r1 = L.act_rec[2 * (size_t)a + 1];
                                    V3 o, d;
                                    float tmin, tmax;
                                    gen_ray_of(L.ds, pixel_from_record(r0, r1), slot,
                                               slot < rs ? slot_light(L.ds, slot) : 0, o, d, tmin, tmax);
                                    ray_begin(r, st, o, d, tmin, tmax);
                                    if (cull) enter_candidates(r, st, cw, L.slot_box);
                                }
                            } else {
                                float4 o = L.ray_o[(size_t)rid * L.ray_stride];
                                float4 d = L.ray_d[(size_t)rid * L.ray_stride];
                                ray_begin(r, st, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), o.w, d.w);
                            }
                        }
                    }
                }
                chunk_next += take;
            }
        }
        if (__ballot(ray != kNone) == 0ull) {
            if (!open) break;
            continue;
        }
        if (ray != kNone) {
            for (;;) {
                int res = ray_step<false, WIDE, StackT<SHORT, DEPTH>, CULLED>(r, st, tnodes, trecs, nullptr,
                                                                          nullptr, tlas_lds, L.max_steps);
                // a ray with nothing left to visit ends in this step, not in one more
                if (res == kStepContinue && r.next == kNone && st.sp == 0) res = kStepDone;
                if (res != kStepContinue) {
                    L.hits[ray] = res == kStepHit ? 1 : 0;
                    if (res == kStepOverflow) atomicOr(&L.counters[kCntError], 1u);
                    ray = kNone;
                    break;
                }
                if (open && 64u - (uint32_t)__popcll(__ballot(true)) >= refill) break;
            }
        }
    }
}

// Thread q zeroes the counts of pixel q if it is background and resolves
// active pixel q (in active order, so a wave's hit reads are contiguous).
__global__ __launch_bounds__(kBlock) void k_resolve(const DevScene* ds, const uint32_t* act_index,
                                                    const uint32_t* act_list, const uint8_t* hits,
                                                    const uint32_t* counters, uint32_t npix,
                                                    uint8_t* shadow, uint8_t* ao) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= npix) return;
    const uint32_t A = counters[kCntActive];
    if (act_index[q] == kNone) resolve_pixel(ds, kNone, A, hits, npix, q, shadow, ao);
    if (q < A) resolve_pixel(ds, q, A, hits, npix, act_list[q], shadow, ao);
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

void launch_cull_scan(hipStream_t st, uint32_t* counts, uint32_t n, uint32_t* counters) {
    k_cull_scan<<<1, 1024, 0, st>>>(counts, n, counters);
}

luzrt_status launch_trace(luzrt_ctx* ctx, const TraceLaunch& L) {
    uint32_t groups = (uint32_t)ctx->num_cus * ctx->groups_per_cu;
    if (groups == 0) groups = 1;
    size_t spill_bytes = (size_t)groups * kBlock * kSpill * 4;
    luzrt_status s = ensure(ctx, ctx->spill, spill_bytes);
    if (s != LUZRT_OK) return s;
    TraceLaunch l = L;
    l.spill = (uint32_t*)ctx->spill.ptr;
    l.max_steps = ctx->max_steps;
    uint32_t refill = ctx->refill_threshold ? ctx->refill_threshold : 64u;
    auto launch = [&](auto kern) { kern<<<groups, kBlock, 0, ctx->stream>>>(l, refill); };
    const uint32_t sd = ctx->stack_entries;
    if (L.nscenes > 0) {
        if (L.wide || L.nscenes > kMaxBatch)
            return fail(ctx, LUZRT_E_UNSUPPORTED, "multi-scene traversal: binary BVHs, <= %u scenes",
                        kMaxBatch);
        if (L.live_rows) {
            if (sd <= 8) launch(k_trace_any<false, 8, kBatch, kSpill, true>);
            else if (sd <= 16) launch(k_trace_any<false, 16, kBatch, kSpill, true>);
            else launch(k_trace_any<false, 32, kBatch, kSpill, true>);
        } else {
            if (sd <= 8) launch(k_trace_any<false, 8, kBatch>);
            else if (sd <= 16) launch(k_trace_any<false, 16, kBatch>);
            else launch(k_trace_any<false, 32, kBatch>);
        }
    } else if (L.gen && L.live_rows) {
        if (L.wide) {
            if (sd <= 8) launch(k_trace_any<true, 8, kGen, kSpill, true>);
            else if (sd <= 16) launch(k_trace_any<true, 16, kGen, kSpill, true>);
            else launch(k_trace_any<true, 32, kGen, kSpill, true>);
        } else {
            if (sd <= 8) launch(k_trace_any<false, 8, kGen, kSpill, true>);
            else if (sd <= 16) launch(k_trace_any<false, 16, kGen, kSpill, true>);
            else launch(k_trace_any<false, 32, kGen, kSpill, true>);
        }
    } else if (L.gen) {
        if (L.wide) {
            if (sd <= 8) launch(k_trace_any<true, 8, kGen>);
            else if (sd <= 16) launch(k_trace_any<true, 16, kGen>);
            else launch(k_trace_any<true, 32, kGen>);
        } else {
            if (sd <= 8) launch(k_trace_any<false, 8, kGen>);
            else if (sd <= 16) launch(k_trace_any<false, 16, kGen>);
            else launch(k_trace_any<false, 32, kGen>);
        }
    } else if (L.wide) {
        if (sd <= 8) launch(k_trace_any<true, 8, kRays>);
        else if (sd <= 16) launch(k_trace_any<true, 16, kRays>);
        else launch(k_trace_any<true, 32, kRays>);
    } else {
        if (sd <= 8) launch(k_trace_any<false, 8, kRays>);
        else if (sd <= 16) launch(k_trace_any<false, 16, kRays>);
        else launch(k_trace_any<false, 32, kRays>);
    }
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

// Counters are read back after the call to size the ray count.
luzrt_status visibility_core(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* g,
                             const luzrt_scene* scene, uint8_t* shadow, uint8_t* ao, bool resolve) {
    DevScene ds;
    luzrt_status s = upload_scene(ctx, scene, &ds, ray_group_for(ctx, tlas, scene), true);
    if (s != LUZRT_OK) return s;
    luzrt_gbuffer gl;
    if ((s = check_gbuffer(ctx, g, &gl)) != LUZRT_OK) return s;
    const uint32_t W = g->width, H = g->height;
    const size_t npix = (size_t)W * H;
    const size_t max_rays = npix * ds.r_total;
    if (max_rays > 0xFFFFFFFFull)
        return fail(ctx, LUZRT_E_LIMIT, "%zu rays exceed the 32-bit ray index", max_rays);
    if ((s = ensure(ctx, ctx->counters, kCounterBytes)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->active, npix * 4)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->act_rec, npix * 32)) != LUZRT_OK) return s;
    // act_index holds the per-pixel active index followed by the per-block counts
    const uint32_t tw = ctx->tile_width, th = 64u / tw;
    const size_t ntiles_px_sz = (size_t)((W + tw - 1) / tw) * ((H + th - 1) / th) * 64;
    if ((s = ensure(ctx, ctx->act_index, npix * 4 + 16 + (ntiles_px_sz / kBlock + 2) * 4)) != LUZRT_OK)
        return s;
    if ((s = ensure(ctx, ctx->hits, max_rays + 16)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    const DevScene* dsd = (const DevScene*)ctx->scene.ptr;
    hipStream_t st = ctx->stream;

    stage_begin(ctx, kStageCompact);
    uint32_t ntiles_px = (uint32_t)ntiles_px_sz;
    uint32_t nblocks = cdiv(ntiles_px, kBlock);
    uint32_t* block_counts = (uint32_t*)ctx->act_index.ptr + ((npix + 3) & ~size_t(3));   // 16-byte aligned
    k_compact_count<<<nblocks, kBlock, 0, st>>>(gl, tw, ntiles_px, block_counts);
    k_compact_scan<<<1, 1024, 0, st>>>(block_counts, nblocks, counters, dsd);
    k_compact_write<<<nblocks, kBlock, 0, st>>>(gl, tw, ntiles_px, dsd,
                                               block_counts, (uint32_t*)ctx->active.ptr,
                                               (uint32_t*)ctx->act_index.ptr, (float4*)ctx->act_rec.ptr);
    stage_end(ctx, kStageCompact);
    LUZ_HIP(ctx, hipGetLastError());

    // rays are generated inside the traversal; the compaction scan set their
    // count (the raygen stage times the culling kernels, when they run).  The
    // compaction reads no TLAS data, so it overlaps the TLAS's pending build.
    if ((s = tlas_acquire(ctx, tlas)) != LUZRT_OK) return s;

    // ray-group culling (grouped order, <= 128 TLAS slots, <= 64 rays per pixel)
    const uint32_t ngroups = (uint32_t)((npix + kCullGroup - 1) / kCullGroup);
    const bool cull = ctx->group_cull && ds.r_total > 0 && ds.r_total <= kMaxCullSlots &&
                      tlas->count > 0 && tlas->count <= kMaxCullInst && ds.ray_group == kCullGroup &&
                      (uint64_t)ngroups * kCullGroup * ds.r_total <= 0xFFFFFFFFull;
    ctx->last_culled = cull;
    const uint32_t kinds = (uint32_t)scene->num_lights + 1;
    if (cull) {
        if ((s = ensure(ctx, ctx->cull_cand, (size_t)ngroups * kinds * kCandWords * 8)) != LUZRT_OK) return s;
        const size_t counts_off = ((size_t)ngroups * 8 + 255) & ~size_t(255);   // 16-byte aligned counts
        if ((s = ensure(ctx, ctx->cull_bits, counts_off + (size_t)ngroups * 4 + 16)) != LUZRT_OK) return s;
        if ((s = ensure(ctx, ctx->cull_rows, (size_t)ngroups * ds.r_total * 4)) != LUZRT_OK) return s;
        uint64_t* bits = (uint64_t*)ctx->cull_bits.ptr;
        uint32_t* counts = (uint32_t*)((uint8_t*)ctx->cull_bits.ptr + counts_off);
        const uint32_t gblocks = cdiv(ngroups, kBlock / 64);
        stage_begin(ctx, kStageRaygen);
        k_cull<<<gblocks, kBlock, 0, st>>>((const float4*)ctx->act_rec.ptr, counters, dsd,
                                           tlas_slot_boxes(tlas), tlas->count, ngroups, kinds,
                                           (uint64_t*)ctx->cull_cand.ptr, bits, counts);
        k_cull_scan<<<1, 1024, 0, st>>>(counts, ngroups, counters);
        k_cull_rows<<<gblocks, kBlock, 0, st>>>(counters, dsd, ngroups, bits, counts,
                                                (uint32_t*)ctx->cull_rows.ptr, (uint8_t*)ctx->hits.ptr);
        stage_end(ctx, kStageRaygen);
        LUZ_HIP(ctx, hipGetLastError());
    }

    stage_begin(ctx, kStageTrace);
    if (ds.r_total > 0 && tlas->count > 0) {
        TraceLaunch L{};
        if (cull) {
            L.live_rows = (const uint32_t*)ctx->cull_rows.ptr;
            L.cand = (const uint64_t*)ctx->cull_cand.ptr;
            L.slot_box = tlas_slot_boxes(tlas);
            L.cand_kinds = kinds;
        }
        L.tlas_nodes = tlas->nodes;
        L.recs = tlas->wide ? tlas->recs_w : tlas->recs;
        L.tlas_nodes_count = tlas->node_count;
        L.recs_count = tlas->count;
        L.hits = (uint8_t*)ctx->hits.ptr;
        L.counters = counters;
        L.wide = tlas->wide;
        L.gen = true;
        L.ds = dsd;
        L.act_rec = (const float4*)ctx->act_rec.ptr;
        if ((s = launch_trace(ctx, L)) != LUZRT_OK) return s;
    } else if (ds.r_total > 0) {
        // empty TLAS: nothing can be hit
        if ((s = zero_fill(ctx, ctx->hits.ptr, max_rays)) != LUZRT_OK) return s;
    }
    stage_end(ctx, kStageTrace);
    if ((s = tlas_release(ctx)) != LUZRT_OK) return s;   // the rest reads no TLAS data

    if (!resolve) return LUZRT_OK;
    stage_begin(ctx, kStageResolve);
    k_resolve<<<cdiv(npix, kBlock), kBlock, 0, st>>>(dsd, (const uint32_t*)ctx->act_index.ptr,
                                                     (const uint32_t*)ctx->active.ptr,
                                                     (const uint8_t*)ctx->hits.ptr, counters,
                                                     (uint32_t)npix, shadow, ao);
    stage_end(ctx, kStageResolve);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

// ------------------------------------------------ visibility without allocation
// luzrt_trace_visibility_static runs the same kernels as luzrt_trace_visibility
// on work buffers carved from one file-scope device array (`scratch`), so the
// call allocates nothing.  Limits: 2^19 pixels, 2^24 rays (pixels x ray slots),
// a traversal of at most kStaticGroups persistent workgroups whose stacks hold
// kStaticDepth entries (16 in LDS, the rest spilled); a deeper ray is reported
// through the error latch like any traversal overflow.
constexpr size_t kStaticPixels = size_t(1) << 19;
constexpr size_t kStaticRays = size_t(1) << 24;
constexpr int kStaticDepth = 48;
constexpr int kStaticShort = 16;
constexpr size_t kStaticSpillBytes = size_t(24) << 20;
constexpr uint32_t kStaticGroups =
    (uint32_t)(kStaticSpillBytes / (kBlock * (kStaticDepth - kStaticShort) * 4));   // 768
constexpr size_t kStaticBlockBytes = size_t(64) << 10;   // compaction block counts
constexpr size_t kOffCounters = 0;
constexpr size_t kOffScene = kCounterBytes;
constexpr size_t kOffActive = 16384;
constexpr size_t kOffIndex = kOffActive + kStaticPixels * 4;
constexpr size_t kOffBlocks = kOffIndex + kStaticPixels * 4;
constexpr size_t kOffRec = kOffBlocks + kStaticBlockBytes;
constexpr size_t kOffHits = kOffRec + kStaticPixels * 32;
constexpr size_t kOffSpill = kOffHits + kStaticRays + 256;
constexpr size_t kStaticBytes = kOffSpill + kStaticSpillBytes;   // 60.1 MiB
static_assert(sizeof(DevScene) <= kOffActive - kOffScene, "scene block fits its slot");
static_assert(kOffScene % 256 == 0, "aligned scene block");
static_assert(kOffRec % 16 == 0 && kOffSpill % 256 == 0, "aligned regions");

// Each GPU that loads the library holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns.
static __device__ __align__(256) uint8_t scratch[kStaticBytes];

void static_scratch_init(luzrt_ctx* ctx) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    ctx->static_scratch = e == hipSuccess ? p : nullptr;
    ctx->static_scratch_error = e;
}

namespace {

// The scene block reaches the scratch as kernel arguments (no host staging
// buffer whose lifetime the asynchronous call would have to outlive).
constexpr uint32_t kSceneChunkWords = 512;
struct SceneChunk {
    uint32_t w[kSceneChunkWords];
};

__global__ void k_put_scene(uint32_t* dst, SceneChunk c, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = c.w[i];
}

// Clears the call's counters (the error latch included: the last call's was
// handed to the context) before the compaction sets them.
__global__ void k_static_init(uint32_t* counters) {
    if (threadIdx.x < kCntCount) counters[threadIdx.x] = 0u;
    clear_claims(counters, threadIdx.x);
}

// Hands the call's ray and pixel counts and its traversal error to the
// context's counters, where luzrt_get_timings and luzrt_sync read them.
__global__ void k_static_latch(const uint32_t* counters, uint32_t* ctx_counters) {
    if (threadIdx.x == 0) {
        ctx_counters[kCntActive] = counters[kCntActive];
        ctx_counters[kCntRays] = counters[kCntRays];
        if (counters[kCntError]) atomicOr(&ctx_counters[kCntError], counters[kCntError]);
    }
}

}  // namespace

}  // namespace luz

using namespace luz;

extern "C" luzrt_status luzrt_trace_rays(luzrt_ctx* ctx, const luzrt_tlas* tlas, const float* rays,
                                         uint32_t count, uint8_t* hits) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || (count && (!rays || !hits)))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_trace_rays: null argument");
    if (((uintptr_t)rays & 15u) != 0)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_trace_rays: rays must be 16-byte aligned");
    if (count == 0) return LUZRT_OK;
    luzrt_status s = ensure(ctx, ctx->counters, kCounterBytes);
    if (s != LUZRT_OK) return s;
    if ((s = tlas_acquire(ctx, tlas)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    stage_begin(ctx, kStageTrace);
    k_vis_init<<<1, 64, 0, ctx->stream>>>(counters, count);
    if (tlas->count == 0) {
        if ((s = zero_fill(ctx, hits, count)) != LUZRT_OK) return s;
    } else {
        TraceLaunch L{};
        L.tlas_nodes = tlas->nodes;
        L.recs = tlas->wide ? tlas->recs_w : tlas->recs;
        L.tlas_nodes_count = tlas->node_count;
        L.recs_count = tlas->count;
        L.ray_o = (const float4*)rays;
        L.ray_d = (const float4*)(rays + 4);
        L.ray_stride = 2;
        L.hits = hits;
        L.counters = counters;
        L.wide = tlas->wide;
        if ((s = launch_trace(ctx, L)) != LUZRT_OK) return s;
    }
    stage_end(ctx, kStageTrace);
    ctx->last_rays = count;
    return LUZRT_OK;
}

extern "C" luzrt_status luzrt_trace_visibility(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                               const luzrt_gbuffer* g, const luzrt_scene* scene,
                                               uint8_t* shadow, uint8_t* ao) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !g || !scene || !ao || (scene->num_lights > 0 && !shadow))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_trace_visibility: null argument");
    return visibility_core(ctx, tlas, g, scene, shadow, ao, true);
}

extern "C" luzrt_status luzrt_trace_visibility_static(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                                      const luzrt_gbuffer* g, const luzrt_scene* scene,
                                                      uint8_t* shadow, uint8_t* ao) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !g || !scene || !ao || (scene->num_lights > 0 && !shadow))
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_trace_visibility_static: null argument");
    if (!ctx->static_scratch)
        return fail(ctx, LUZRT_E_DEVICE, "luzrt_trace_visibility_static: no scratch on this device (%s)",
                    hipGetErrorString(ctx->static_scratch_error));
    DevScene ds;
    luzrt_status s = fill_scene(ctx, scene, &ds, true);
    if (s != LUZRT_OK) return s;
    ds.ray_group = ray_group_for(ctx, tlas);
    luzrt_gbuffer gl;
    if ((s = check_gbuffer(ctx, g, &gl)) != LUZRT_OK) return s;
    const uint32_t W = g->width, H = g->height;
    const size_t npix = (size_t)W * H;
    const uint32_t tw = ctx->tile_width, th = 64u / tw;
    const size_t ntiles_px_sz = (size_t)((W + tw - 1) / tw) * ((H + th - 1) / th) * 64;
    const uint32_t nblocks = cdiv(ntiles_px_sz, kBlock);
    if (npix > kStaticPixels)
        return fail(ctx, LUZRT_E_LIMIT, "luzrt_trace_visibility_static: %ux%u exceeds %zu pixels", W, H,
                    kStaticPixels);
    if (npix * ds.r_total > kStaticRays)
        return fail(ctx, LUZRT_E_LIMIT,
                    "luzrt_trace_visibility_static: %zu pixels x %u ray slots exceed %zu rays", npix,
                    ds.r_total, kStaticRays);
    if ((size_t)(nblocks + 2) * 4 > kStaticBlockBytes)
        return fail(ctx, LUZRT_E_LIMIT, "luzrt_trace_visibility_static: %u compaction blocks", nblocks);
    ctx->last_culled = false;   // this path traces every ray
    uint8_t* base = (uint8_t*)ctx->static_scratch;
    uint32_t* counters = (uint32_t*)(base + kOffCounters);
    DevScene* dsd = (DevScene*)(base + kOffScene);
    uint32_t* act_list = (uint32_t*)(base + kOffActive);
    uint32_t* act_index = (uint32_t*)(base + kOffIndex);
    uint32_t* block_counts = (uint32_t*)(base + kOffBlocks);
    float4* act_rec = (float4*)(base + kOffRec);
    uint8_t* hits = base + kOffHits;
    uint32_t* spill = (uint32_t*)(base + kOffSpill);
    hipStream_t st = ctx->stream;

    const uint32_t words = (uint32_t)((sizeof(DevScene) + 3) / 4);
    for (uint32_t w0 = 0; w0 < words; w0 += kSceneChunkWords) {
        SceneChunk c;
        const uint32_t n = words - w0 < kSceneChunkWords ? words - w0 : kSceneChunkWords;
        std::memset(&c, 0, sizeof(c));
        std::memcpy(c.w, reinterpret_cast<const uint8_t*>(&ds) + (size_t)w0 * 4,
                    std::min<size_t>((size_t)n * 4, sizeof(DevScene) - (size_t)w0 * 4));
        k_put_scene<<<1, 256, 0, st>>>(reinterpret_cast<uint32_t*>(dsd) + w0, c, n);
    }
    k_static_init<<<1, 64, 0, st>>>(counters);
    stage_begin(ctx, kStageCompact);
    const uint32_t ntiles_px = (uint32_t)ntiles_px_sz;
    k_compact_count<<<nblocks, kBlock, 0, st>>>(gl, tw, ntiles_px, block_counts);
    k_compact_scan<<<1, 1024, 0, st>>>(block_counts, nblocks, counters, dsd);
    k_compact_write<<<nblocks, kBlock, 0, st>>>(gl, tw, ntiles_px, dsd, block_counts, act_list, act_index,
                                               act_rec);
    stage_end(ctx, kStageCompact);
    LUZ_HIP(ctx, hipGetLastError());

    if ((s = tlas_acquire(ctx, tlas)) != LUZRT_OK) return s;
    stage_begin(ctx, kStageTrace);
    if (ds.r_total > 0 && tlas->count > 0) {
        TraceLaunch L{};
        L.tlas_nodes = tlas->nodes;
        L.recs = tlas->wide ? tlas->recs_w : tlas->recs;
        L.tlas_nodes_count = tlas->node_count;
        L.recs_count = tlas->count;
        L.hits = hits;
        L.counters = counters;
        L.spill = spill;
        L.max_steps = ctx->max_steps;
        L.wide = tlas->wide;
        L.gen = true;
        L.ds = dsd;
        L.act_rec = act_rec;
        uint32_t groups = (uint32_t)ctx->num_cus * ctx->groups_per_cu;
        groups = groups == 0 ? 1u : (groups > kStaticGroups ? kStaticGroups : groups);
        const uint32_t refill = ctx->refill_threshold ? ctx->refill_threshold : 64u;
        if (L.wide)
            k_trace_any<true, kStaticShort, kGen, kStaticDepth><<<groups, kBlock, 0, st>>>(L, refill);
        else
            k_trace_any<false, kStaticShort, kGen, kStaticDepth><<<groups, kBlock, 0, st>>>(L, refill);
        LUZ_HIP(ctx, hipGetLastError());
    } else if (ds.r_total > 0) {
        if ((s = zero_fill(ctx, hits, npix * ds.r_total)) != LUZRT_OK) return s;
    }
    stage_end(ctx, kStageTrace);

    stage_begin(ctx, kStageResolve);
    k_resolve<<<cdiv(npix, kBlock), kBlock, 0, st>>>(dsd, act_index, act_list, hits, counters,
                                                     (uint32_t)npix, shadow, ao);
    k_static_latch<<<1, 64, 0, st>>>(counters, (uint32_t*)ctx->counters.ptr);
    stage_end(ctx, kStageResolve);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}
