// trace.hip — shadow and ambient-occlusion visibility on gfx950.
//
// Replaces the ray queries of light.frag (TraceShadowRay :86-109,
// TraceAORays :111-135, EvaluateShadow :137-146, AO origin :229-231) with a
// wavefront pipeline over coalesced structure-of-arrays buffers:
//   k_vis_init  zero the call's device counters (graph-safe, no memset);
//   k_compact_* non-background pixels in 8x8-tile order -> active list
//               (count, scan, write: deterministic, no atomics);
//   k_raygen    one thread per active pixel writes its shadow and AO rays,
//               slot-major (ray j*A + a), so a wave holds 64 neighbouring
//               pixels of the same light sample or AO sample;
//   k_trace_any persistent any-hit traversal: a wave refills its idle lanes
//               with one atomic when at least `refill` lanes are idle
//               (ballot-based compaction of the wave's work);
//   k_resolve   per pixel, occluded shadow samples per light and unoccluded
//               AO samples, written as u8 planes.
#include "internal.hpp"
#include "traverse.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

__global__ void k_vis_init(uint32_t* counters, uint32_t ray_count) {
    if (threadIdx.x <= kCntRays) counters[threadIdx.x] = threadIdx.x == kCntRays ? ray_count : 0u;
}

__device__ inline void tile_pixel(uint32_t g, uint32_t W, uint32_t& x, uint32_t& y) {
    uint32_t tile = g >> 6, t = g & 63u;
    uint32_t tiles_x = (W + 7) >> 3;
    x = (tile % tiles_x) * 8 + (t & 7u);
    y = (tile / tiles_x) * 8 + (t >> 3);
}

// Compaction of non-background pixels without atomics, in three passes over
// 8x8-tile order: per-block counts, one exclusive scan, per-block writes.  The
// active list is therefore deterministic (tile order) and contiguous per tile.
__device__ inline bool pixel_active(const float* normal, uint32_t g, uint32_t W, uint32_t H,
                                    uint32_t ntiles_px, uint32_t& p) {
    uint32_t x = 0, y = 0;
    p = 0;
    if (g >= ntiles_px) return false;
    tile_pixel(g, W, x, y);
    if (x >= W || y >= H) return false;
    p = y * W + x;
    const float* n = normal + (size_t)p * 3;
    return vlen(v3(n[0], n[1], n[2])) != 0.0f;
}

__device__ inline uint32_t block_prefix(bool active, uint32_t* wave_tot, uint32_t& block_total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t m = __ballot(active);
    uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        uint32_t c = wave_tot[w];
        if (w < wave) off += c;
        tot += c;
    }
    block_total = tot;
    return off + (uint32_t)__popcll(m & lt);
}

__global__ __launch_bounds__(kBlock) void k_compact_count(const float* normal, uint32_t W, uint32_t H,
                                                          uint32_t ntiles_px, uint32_t* block_counts) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    uint32_t p;
    bool active = pixel_active(normal, blockIdx.x * kBlock + threadIdx.x, W, H, ntiles_px, p);
    uint32_t total;
    block_prefix(active, wave_tot, total);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = total;
}

// Exclusive scan of the block counts (one workgroup); writes the total.
__global__ __launch_bounds__(1024) void k_compact_scan(uint32_t* block_counts, uint32_t nblocks,
                                                       uint32_t* counters) {
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nblocks; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nblocks ? block_counts[i] : 0u;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t off = 1; off < 1024; off <<= 1) {
            uint32_t add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        uint32_t incl = sh[threadIdx.x];
        uint32_t c = carry;
        if (i < nblocks) block_counts[i] = c + incl - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) counters[kCntActive] = carry;
}

__global__ __launch_bounds__(kBlock) void k_compact_write(const float* normal, uint32_t W, uint32_t H,
                                                          uint32_t ntiles_px,
                                                          const uint32_t* block_offsets,
                                                          uint32_t* act_list, uint32_t* act_index) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    uint32_t p;
    bool active = pixel_active(normal, g, W, H, ntiles_px, p);
    uint32_t total;
    uint32_t a = block_offsets[blockIdx.x] + block_prefix(active, wave_tot, total);
    bool inside = false;
    if (g < ntiles_px) {
        uint32_t x, y;
        tile_pixel(g, W, x, y);
        inside = x < W && y < H;
    }
    if (active) act_list[a] = p;
    if (inside) act_index[p] = active ? a : kNone;
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

__global__ __launch_bounds__(kBlock) void k_raygen(const float* position, const float* normal,
                                                   uint32_t W, const DevScene* ds,
                                                   const uint32_t* act_list,
                                                   const uint32_t* counters, float4* ray_o,
                                                   float4* ray_d) {
    const uint32_t A = counters[kCntActive];
    uint32_t a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= A) return;
    const luzrt_scene& S = ds->s;
    uint32_t p = act_list[a];
    uint32_t x = p % W, y = p / W;
    V3 P = v3(position[(size_t)p * 3], position[(size_t)p * 3 + 1], position[(size_t)p * 3 + 2]);
    V3 N = v3(normal[(size_t)p * 3], normal[(size_t)p * 3 + 1], normal[(size_t)p * 3 + 2]);
    V3 cam = v3(S.cam_pos[0], S.cam_pos[1], S.cam_pos[2]);
    float lc = vlen(vsub(P, cam));
    float sb = fmax_(lc * 0.01f, 0.05f);
    V3 so = vadd(P, vmul(N, sb));
    const uint32_t frame = (uint32_t)S.frame;
    for (int l = 0; l < S.num_lights; ++l) {
        const luzrt_light& L = S.lights[l];
        int n = L.num_shadow_samples;
        if (n <= 0) continue;
        ShadowSetup ss = shadow_setup(L, P);
        uint32_t j0 = ds->slot_base[l];
        for (int s = 0; s < n; ++s) {
            float u, v;
            sample2(x, y, frame, (uint32_t)(l * 256 + s), u, v);
            float pr = L.radius * sqrtf(u);
            float sn, cs;
            sincos_turns(v, sn, cs);
            float dx = pr * cs, dy = pr * sn;
            V3 d = v3(ss.Lvec.x + dx * ss.T.x + dy * ss.B.x,
                      ss.Lvec.y + dx * ss.T.y + dy * ss.B.y,
                      ss.Lvec.z + dx * ss.T.z + dy * ss.B.z);
            d = vnorm(d);
            size_t idx = (size_t)(j0 + s) * A + a;
            ray_o[idx] = make_float4(so.x, so.y, so.z, 0.001f);
            ray_d[idx] = make_float4(d.x, d.y, d.z, ss.tmax);
        }
    }
    int na = S.ao_num_samples;
    if (na > 0) {
        float ab = lc * 0.01f;
        V3 ao = vadd(P, vmul(N, ab));
        V3 T = fabsf(N.z) > 0.5f ? v3(0.0f, -N.z, N.y) : v3(-N.y, N.x, 0.0f);
        V3 B = vcross(N, T);
        uint32_t j0 = ds->r_shadow;
        for (int s = 0; s < na; ++s) {
            float u, v;
            sample2(x, y, frame, (uint32_t)(16384 + s), u, v);
            float r = sqrtf(u);
            float sn, cs;
            sincos_turns(v * kHemiTurns, sn, cs);
            float hx = r * cs, hy = r * sn;
            float om = 1.0f - u;
            float hz = sqrtf(om > 0.0f ? om : 0.0f);
            V3 d = v3(T.x * hx + B.x * hy + N.x * hz,
                      T.y * hx + B.y * hy + N.y * hz,
                      T.z * hx + B.z * hy + N.z * hz);
            size_t idx = (size_t)(j0 + s) * A + a;
            ray_o[idx] = make_float4(ao.x, ao.y, ao.z, S.ao_min);
            ray_d[idx] = make_float4(d.x, d.y, d.z, S.ao_max);
        }
    }
    if (a == 0) {
        // the trace kernel reads the ray count from here
        const_cast<uint32_t*>(counters)[kCntRays] = A * ds->r_total;
    }
}

// Rays are claimed in chunks of kChunk per wave (one device-scope atomic per
// chunk): a single counter hit by every refill serialises across the 8 XCDs.
constexpr uint32_t kChunk = 2048;

template <bool WIDE>
__global__ __launch_bounds__(kBlock) void k_trace_any(TraceLaunch L, uint32_t refill) {
    __shared__ uint32_t lds[kShort * kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint32_t total = gridDim.x * kBlock;
    Stack st;
    st.ls = lds + tid;
    st.lds_stride = kBlock;
    st.gs = L.spill + (size_t)blockIdx.x * kBlock + tid;
    st.g_stride = total;
    st.sp = 0;
    st.lo = 0;
    const uint32_t nrays = L.counters[kCntRays];
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    RayState r;
    uint32_t ray = kNone;
    bool open = true;
    uint32_t chunk_next = 0, chunk_end = 0;  // wave-uniform
    for (;;) {
        if (open) {
            uint64_t idle = __ballot(ray == kNone);
            uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= refill || nidle == 64u) {
                if (chunk_next >= chunk_end) {
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&L.counters[kCntFetch], kChunk);
                    base = __shfl(base, 0);
                    chunk_next = base;
                    chunk_end = base + kChunk < nrays ? base + kChunk : nrays;
                    if (base >= nrays) open = false;
                }
                uint32_t avail = chunk_end > chunk_next ? chunk_end - chunk_next : 0u;
                uint32_t take = nidle < avail ? nidle : avail;
                if (ray == kNone) {
                    uint32_t rank = (uint32_t)__popcll(idle & lt);
                    if (rank < take) {
                        uint32_t rid = chunk_next + rank;
                        ray = rid;
                        float4 o = L.ray_o[(size_t)rid * L.ray_stride];
                        float4 d = L.ray_d[(size_t)rid * L.ray_stride];
                        ray_begin(r, st, v3(o.x, o.y, o.z), v3(d.x, d.y, d.z), o.w, d.w);
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
                int res = ray_step<false, WIDE>(r, st, L.tlas_nodes, L.recs, nullptr, nullptr);
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

__global__ __launch_bounds__(kBlock) void k_resolve(const DevScene* ds, const uint32_t* act_index,
                                                    const uint8_t* hits, const uint32_t* counters,
                                                    uint32_t npix, uint8_t* shadow, uint8_t* ao) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    const uint32_t A = counters[kCntActive];
    const luzrt_scene& S = ds->s;
    uint32_t a = act_index[p];
    for (int l = 0; l < S.num_lights; ++l) {
        uint32_t cnt = 0;
        int n = S.lights[l].num_shadow_samples;
        if (a != kNone && n > 0) {
            uint32_t j0 = ds->slot_base[l];
            for (int s = 0; s < n; ++s) cnt += hits[(size_t)(j0 + s) * A + a];
        }
        shadow[(size_t)l * npix + p] = (uint8_t)cnt;
    }
    uint32_t miss = 0;
    if (a != kNone) {
        uint32_t j0 = ds->r_shadow;
        for (int s = 0; s < S.ao_num_samples; ++s) miss += 1u - hits[(size_t)(j0 + s) * A + a];
    }
    ao[p] = (uint8_t)miss;
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

luzrt_status launch_trace(luzrt_ctx* ctx, const TraceLaunch& L) {
    uint32_t groups = (uint32_t)ctx->num_cus * ctx->groups_per_cu;
    if (groups == 0) groups = 1;
    size_t spill_bytes = (size_t)groups * kBlock * kSpill * 4;
    luzrt_status s = ensure(ctx, ctx->spill, spill_bytes);
    if (s != LUZRT_OK) return s;
    TraceLaunch l = L;
    l.spill = (uint32_t*)ctx->spill.ptr;
    uint32_t refill = ctx->refill_threshold ? ctx->refill_threshold : 64u;
    if (L.wide)
        k_trace_any<true><<<groups, kBlock, 0, ctx->stream>>>(l, refill);
    else
        k_trace_any<false><<<groups, kBlock, 0, ctx->stream>>>(l, refill);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

// Counters are read back after the call to size the ray count.
static luzrt_status visibility_core(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* g,
                                    const luzrt_scene* scene, uint8_t* shadow, uint8_t* ao) {
    DevScene ds;
    luzrt_status s = upload_scene(ctx, scene, &ds);
    if (s != LUZRT_OK) return s;
    if ((s = check_gbuffer(ctx, g)) != LUZRT_OK) return s;
    const uint32_t W = g->width, H = g->height;
    const size_t npix = (size_t)W * H;
    const size_t max_rays = npix * ds.r_total;
    if (max_rays > 0xFFFFFFFFull)
        return fail(ctx, LUZRT_E_LIMIT, "%zu rays exceed the 32-bit ray index", max_rays);
    if ((s = ensure(ctx, ctx->counters, 64)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->active, npix * 4)) != LUZRT_OK) return s;
    // act_index holds the per-pixel active index followed by the per-block counts
    const size_t ntiles_px_sz = (size_t)((W + 7) / 8) * ((H + 7) / 8) * 64;
    if ((s = ensure(ctx, ctx->act_index, npix * 4 + (ntiles_px_sz / kBlock + 2) * 4)) != LUZRT_OK)
        return s;
    if ((s = ensure(ctx, ctx->rays_o, max_rays * 16 + 16)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->rays_d, max_rays * 16 + 16)) != LUZRT_OK) return s;
    if ((s = ensure(ctx, ctx->hits, max_rays + 16)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    const DevScene* dsd = (const DevScene*)ctx->scene.ptr;
    hipStream_t st = ctx->stream;

    stage_begin(ctx, kStageCompact);
    k_vis_init<<<1, 64, 0, st>>>(counters, 0u);
    uint32_t ntiles_px = ((W + 7) / 8) * ((H + 7) / 8) * 64;
    uint32_t nblocks = cdiv(ntiles_px, kBlock);
    uint32_t* block_counts = (uint32_t*)ctx->act_index.ptr + npix;
    k_compact_count<<<nblocks, kBlock, 0, st>>>(g->normal, W, H, ntiles_px, block_counts);
    k_compact_scan<<<1, 1024, 0, st>>>(block_counts, nblocks, counters);
    k_compact_write<<<nblocks, kBlock, 0, st>>>(g->normal, W, H, ntiles_px, block_counts,
                                               (uint32_t*)ctx->active.ptr,
                                               (uint32_t*)ctx->act_index.ptr);
    stage_end(ctx, kStageCompact);
    LUZ_HIP(ctx, hipGetLastError());

    stage_begin(ctx, kStageRaygen);
    if (ds.r_total > 0)
        k_raygen<<<cdiv(npix, kBlock), kBlock, 0, st>>>(g->position, g->normal, W, dsd,
                                                        (const uint32_t*)ctx->active.ptr, counters,
                                                        (float4*)ctx->rays_o.ptr,
                                                        (float4*)ctx->rays_d.ptr);
    stage_end(ctx, kStageRaygen);
    LUZ_HIP(ctx, hipGetLastError());

    stage_begin(ctx, kStageTrace);
    if (ds.r_total > 0 && tlas->count > 0) {
        TraceLaunch L{};
        L.tlas_nodes = tlas->nodes;
        L.recs = tlas->wide ? tlas->recs_w : tlas->recs;
        L.ray_o = (const float4*)ctx->rays_o.ptr;
        L.ray_d = (const float4*)ctx->rays_d.ptr;
        L.ray_stride = 1;
        L.hits = (uint8_t*)ctx->hits.ptr;
        L.counters = counters;
        L.wide = tlas->wide;
        if ((s = launch_trace(ctx, L)) != LUZRT_OK) return s;
    } else if (ds.r_total > 0) {
        // empty TLAS: nothing can be hit
        LUZ_HIP(ctx, hipMemsetAsync(ctx->hits.ptr, 0, max_rays, st));
    }
    stage_end(ctx, kStageTrace);

    stage_begin(ctx, kStageResolve);
    k_resolve<<<cdiv(npix, kBlock), kBlock, 0, st>>>(dsd, (const uint32_t*)ctx->act_index.ptr,
                                                     (const uint8_t*)ctx->hits.ptr, counters,
                                                     (uint32_t)npix, shadow, ao);
    stage_end(ctx, kStageResolve);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

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
    luzrt_status s = ensure(ctx, ctx->counters, 64);
    if (s != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    stage_begin(ctx, kStageTrace);
    k_vis_init<<<1, 64, 0, ctx->stream>>>(counters, count);
    if (tlas->count == 0) {
        LUZ_HIP(ctx, hipMemsetAsync(hits, 0, count, ctx->stream));
    } else {
        TraceLaunch L{};
        L.tlas_nodes = tlas->nodes;
        L.recs = tlas->wide ? tlas->recs_w : tlas->recs;
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
    return visibility_core(ctx, tlas, g, scene, shadow, ao);
}
