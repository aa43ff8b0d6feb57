// batch.hip — many independent scenes rendered as one frame on gfx950.
//
// A serving frame over S scenes of uneven size and resolution is a fixed
// sequence of launches, each covering every scene at once:
//   staging copy      instances and scene blocks of all scenes (one H2D copy)
//   k_refit_*_batched every registered BLAS refit from its vertices
//   k_tlas_build_batched  one workgroup per scene rebuilds its TLAS
//   k_bprimary        persistent closest-hit G-buffer of every scene with a camera
//   k_bcompact_*      per-scene non-background pixel lists (flat grid of tile blocks)
//   k_bray_base       per-scene ray ranges in one global ray index space
//   k_trace_any<..., kBatch>  ONE persistent traversal over the rays of all
//                     scenes, each generated at claim time from its scene's
//                     G-buffer (ray_row order per scene)
//   k_bresolve, k_bshade     per-pixel counts and FP32 GGX radiance
// Flat grids find their scene from per-scene block bases (binary search on
// values that are uniform across the workgroup).  The sequence is captured
// once into a hipGraph and replayed; it is recaptured when instance counts
// or the trace configuration change.  Nothing in it is a memset node.
//
// Per scene the arithmetic is that of the single-scene calls (lightpath.hpp),
// so a batched frame reproduces luzrt_primary_gbuffer + luzrt_light_pass of
// each scene bit for bit.
#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "cull.hpp"
#include "internal.hpp"
#include "lightpath.hpp"

namespace luz {

struct BatchSceneDev {
    luzrt_gbuffer g;
    luzrt_gbuffer_out gout;
    luzrt_camera cam;
    const DevScene* ds;
    const luzrt_material* mats;
    uint8_t* shadow;
    uint8_t* ao;
    float4* out;
    uint32_t* act_list;
    uint32_t* act_index;
    uint32_t* block_counts;
    const Node* tlas_nodes;    // null: empty TLAS
    const InstRec* recs;
    const InstAux* aux;
    uint32_t nmats;
    uint32_t npix;
    uint32_t ntiles_px;        // tile-ordered pixel slots (compaction)
    uint32_t nblocks;          // compaction blocks
    uint32_t has_camera;
    uint32_t pad[3];
};

namespace {

constexpr int kBlock = 256;

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

// Last s < n with base[s] <= v (base[0] = 0).
__device__ inline uint32_t find_scene(const uint32_t* base, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// Device-side tables of a batch (one allocation).
struct BatchTables {
    const BatchSceneDev* sc;
    const uint32_t* cblk;   // [S+1] first compaction block of each scene
    const uint32_t* pblk;   // [S+1] first pixel block of each scene
    const uint32_t* ppx;    // [S+1] first primary tile-pixel of each scene (cameras only)
    uint32_t* active;       // [S] active pixels of the frame
    uint32_t* ray_base;     // [S+1] first ray of each scene; [S] = total
    uint32_t* gbase;        // [S+1] first 64-pixel culling group of each scene; [S] = total
    uint32_t n;
};

// Per-frame counters; the build-error flag and the traversal error latch
// kCntError stay set until luzrt_batch_rays reports (and clears) them.
__global__ void k_batch_init(uint32_t* counters) {
    if (threadIdx.x < kCntCount && threadIdx.x != kCntBuildError && threadIdx.x != kCntError)
        counters[threadIdx.x] = 0u;
    clear_claims(counters, threadIdx.x);
}

// Persistent closest-hit primary rays over the 8x8-tile pixel slots of every
// scene with a camera, in one flat index space; each wave takes the next tile
// (lane = pixel of the tile) from counters[kCntPrimary] until none is left.
__global__ __launch_bounds__(kBlock) void k_bprimary(BatchTables T, uint32_t* spill,
                                                     uint32_t* counters, uint32_t chunk) {
    __shared__ uint32_t lds[kShort * kBlock];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t total = gridDim.x * kBlock;
    Stack st;
    st.init(lds + tid, kBlock, spill + (size_t)blockIdx.x * kBlock + tid, total);
    const uint32_t ntiles = T.ppx[T.n] >> 6;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&counters[kCntPrimary], chunk);
        base = __shfl(base, 0);
        if (base >= ntiles) break;
        const uint32_t end = min(base + chunk, ntiles);
        for (uint32_t tile = base; tile < end; ++tile) {
            const uint32_t gg = tile * 64u;
            uint32_t s = find_scene(T.ppx, T.n, gg);
            const BatchSceneDev& S = T.sc[s];
            uint32_t g = gg - T.ppx[s];
            const uint32_t W = S.gout.width, H = S.gout.height;
            const uint32_t tiles_x = (W + 7) / 8;
            uint32_t t = g >> 6;
            uint32_t x = (t % tiles_x) * 8 + (lane & 7u), y = (t / tiles_x) * 8 + (lane >> 3);
            if (x < W && y < H)
                primary_pixel(S.tlas_nodes, S.recs, S.aux, S.cam, S.mats, S.nmats, S.gout, x, y,
                              st, &counters[kCntError], S.tlas_nodes == nullptr);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_bcompact_count(BatchTables T, uint32_t tw) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    const uint32_t s = find_scene(T.cblk, T.n, blockIdx.x);
    const BatchSceneDev& S = T.sc[s];
    const uint32_t b = blockIdx.x - T.cblk[s];
    uint32_t p;
    bool inside;
    bool active = pixel_active(S.g.normal, b * kBlock + threadIdx.x, S.g.width, S.g.height, tw,
                               S.ntiles_px, p, inside);
    uint32_t total;
    block_prefix(active, wave_tot, total);
    if (threadIdx.x == 0) S.block_counts[b] = total;
}

// One workgroup per scene: exclusive scan of its block counts.
__global__ __launch_bounds__(1024) void k_bcompact_scan(BatchTables T) {
    __shared__ uint32_t sh[1024];
    __shared__ uint32_t carry;
    const BatchSceneDev& S = T.sc[blockIdx.x];
    uint32_t* block_counts = S.block_counts;
    const uint32_t nblocks = S.nblocks;
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
    if (threadIdx.x == 0) T.active[blockIdx.x] = carry;
}

// Ray ranges: scene s owns rays [ray_base[s], ray_base[s] + active[s] * r_total[s]).
// The host has checked that the sum over all pixels fits 32 bits.
// Culling groups: scene s owns groups [gbase[s], gbase[s + 1]), ceil(active[s] / 64).
// A culled frame starts every scene's rays at a multiple of `align` (16: the
// culled rows' hit bytes are then written as 16-byte vectors); kCntRays is
// the number of rays either way.
__global__ void k_bray_base(BatchTables T, uint32_t* counters, uint32_t align) {
    if (threadIdx.x != 0) return;
    uint32_t base = 0, gb = 0, rays = 0;
    for (uint32_t s = 0; s < T.n; ++s) {
        T.ray_base[s] = base;
        T.gbase[s] = gb;
        const uint32_t r = T.active[s] * T.sc[s].ds->r_total;
        rays += r;
        base = (base + r + align - 1) / align * align;
        gb += (T.active[s] + kCullGroup - 1) / kCullGroup;
    }
    T.ray_base[T.n] = base;
    T.gbase[T.n] = gb;
    counters[kCntRays] = rays;
}

// Ray-group culling of a batched frame (cull.hpp): one wave per 64-pixel group
// of the batch's group space (groups past the frame's total get no live row);
// the group's pixel frames come from its scene's G-buffer planes as the
// traversal generates them.
__global__ __launch_bounds__(kBlock) void k_bcull(BatchTables T, const BatchTlas* tl, uint32_t groups_max,
                                                  uint32_t kinds, uint64_t* cand, uint64_t* live_bits,
                                                  uint32_t* live_count) {
    __shared__ uint64_t s_mask[kBlock / 64][(LUZRT_MAX_LIGHTS + 1) * kCandWords];
    __shared__ luzrt_light s_light[LUZRT_MAX_LIGHTS];
    __shared__ uint32_t s_slot[LUZRT_MAX_LIGHTS];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (kBlock / 64) + wave;
    // when the workgroup's four groups are of one scene (all but a few), its
    // lights and slot bases are staged in LDS as in k_cull
    const uint32_t total = T.gbase[T.n];
    const uint32_t gf = blockIdx.x * (kBlock / 64);
    const uint32_t gl = min(gf + kBlock / 64, total) - 1;
    const bool one_scene = gf < total && find_scene(T.gbase, T.n, gf) == find_scene(T.gbase, T.n, gl);
    if (one_scene) {
        const DevScene* ds0 = tl[find_scene(T.gbase, T.n, gf)].ds;
        const uint32_t nlw = (uint32_t)ds0->s.num_lights * (uint32_t)(sizeof(luzrt_light) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(ds0->s.lights);
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_light);
        for (uint32_t i = threadIdx.x; i < nlw; i += kBlock) dst[i] = src[i];
        for (uint32_t i = threadIdx.x; i < (uint32_t)ds0->s.num_lights; i += kBlock) s_slot[i] = ds0->slot_base[i];
        __syncthreads();
    }
    if (g >= groups_max) return;
    if (g >= total) {
        if (lane == 0) {
            live_bits[g] = 0;
            live_count[g] = 0;
        }
        return;
    }
    const uint32_t s = find_scene(T.gbase, T.n, g);
    const BatchTlas& bt = tl[s];
    const DevScene* __restrict__ ds = bt.ds;
    const uint32_t A = T.active[s], g0 = (g - T.gbase[s]) * kCullGroup;
    const uint32_t a = g0 + lane;
    // lanes past the last active pixel repeat the group's first pixel
    const uint32_t ar = a < A ? a : g0;
    const PixelFrame f = pixel_frame(bt.position, bt.normal, bt.width, ds->s, bt.act_list[ar]);
    uint64_t bits;
    if (one_scene)
        bits = cull_group(f, s_light, s_slot, (uint32_t)ds->s.num_lights, ds->r_shadow, ds->r_total,
                          ds->s.ao_max, bt.slot_box, bt.ninst, s_mask[wave], cand + (size_t)g * kinds * kCandWords);
    else
        bits = cull_group(f, ds->s.lights, ds->slot_base, (uint32_t)ds->s.num_lights, ds->r_shadow, ds->r_total,
                          ds->s.ao_max, bt.slot_box, bt.ninst, s_mask[wave], cand + (size_t)g * kinds * kCandWords);
    if (lane == 0) {
        live_bits[g] = bits;
        live_count[g] = (uint32_t)__popcll(bits);
    }
}

// Live rows of the batch's groups (codes g << 6 | j) and the culled rows' hit
// bytes (misses), as k_cull_rows.
__global__ __launch_bounds__(kBlock) void k_bcull_rows(BatchTables T, uint32_t groups_max,
                                                       const uint64_t* live_bits, const uint32_t* row_base,
                                                       uint32_t* rows, uint8_t* hits) {
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t g = blockIdx.x * (kBlock / 64) + wave;
    if (g >= groups_max || g >= T.gbase[T.n]) return;
    const uint32_t s = find_scene(T.gbase, T.n, g);
    const uint32_t A = T.active[s], g0 = (g - T.gbase[s]) * kCullGroup;
    const uint32_t R = T.sc[s].ds->r_total;
    const uint32_t w = A - g0 < kCullGroup ? A - g0 : kCullGroup;
    cull_rows_lane(live_bits[g], R, w, g << 6, rows + row_base[g], hits + T.ray_base[s] + (size_t)g0 * R);
}

__global__ __launch_bounds__(kBlock) void k_bcompact_write(BatchTables T, uint32_t tw) {
    __shared__ uint32_t wave_tot[kBlock / 64];
    const uint32_t s = find_scene(T.cblk, T.n, blockIdx.x);
    const BatchSceneDev& S = T.sc[s];
    const uint32_t b = blockIdx.x - T.cblk[s];
    uint32_t p;
    bool inside;
    bool active = pixel_active(S.g.normal, b * kBlock + threadIdx.x, S.g.width, S.g.height, tw,
                               S.ntiles_px, p, inside);
    uint32_t total;
    uint32_t a = S.block_counts[b] + block_prefix(active, wave_tot, total);
    if (active) S.act_list[a] = p;
    if (inside) S.act_index[p] = active ? a : kNone;
}

// As k_resolve: thread q of scene s zeroes background pixel q and resolves
// active pixel q.
__global__ __launch_bounds__(kBlock) void k_bresolve(BatchTables T, const uint8_t* hits) {
    const uint32_t s = find_scene(T.pblk, T.n, blockIdx.x);
    const BatchSceneDev& S = T.sc[s];
    const uint32_t q = (blockIdx.x - T.pblk[s]) * kBlock + threadIdx.x;
    if (q >= S.npix) return;
    const uint32_t A = T.active[s];
    const uint8_t* h = hits + T.ray_base[s];
    if (S.act_index[q] == kNone) resolve_pixel(S.ds, kNone, A, h, S.npix, q, S.shadow, S.ao);
    if (q < A) resolve_pixel(S.ds, q, A, h, S.npix, S.act_list[q], S.shadow, S.ao);
}

// The scene block is read through a __restrict__ pointer (nothing the kernel
// writes aliases it), so its lights come through scalar loads; the sRGB table
// is staged in LDS as in the single-scene light pass.
__global__ __launch_bounds__(kBlock) void k_bshade(BatchTables T, const float* __restrict__ lut) {
    __shared__ float s_lut[256];
    s_lut[threadIdx.x] = lut[threadIdx.x];
    __syncthreads();
    const uint32_t s = find_scene(T.pblk, T.n, blockIdx.x);
    const BatchSceneDev& S = T.sc[s];
    const uint32_t p = (blockIdx.x - T.pblk[s]) * kBlock + threadIdx.x;
    if (p >= S.npix) return;
    const DevScene* __restrict__ ds = S.ds;
    S.out[p] = shade_pixel(S.g, ds, s_lut, CountVisibility{S.shadow, S.ao, S.npix}, p);
}

}  // namespace
}  // namespace luz

using namespace luz;

struct luzrt_batch {
    uint32_t n = 0;
    std::vector<luzrt_batch_scene> desc;
    std::vector<BatchSceneDev> sc;      // host copy of the device scene table
    std::vector<luzrt_camera> cams;
    std::vector<uint32_t> has_cam;
    std::vector<std::vector<luzrt_material>> mats;
    // pinned staging: [DevScene x n][InstIn x sum(capacity)], copied once per frame
    uint8_t* host_stage = nullptr;
    uint8_t* dev_stage = nullptr;
    size_t stage_bytes = 0;
    std::vector<size_t> inst_off;       // byte offset of each scene's instances
    std::vector<uint32_t> inst_n;       // staged instance counts (after filtering)
    std::vector<uint32_t> graph_n;      // counts the graph was captured with
    std::vector<uint8_t> staged;        // scene has staged instances
    // per-scene device work arrays and tables
    uint8_t* dev_work = nullptr;
    uint8_t* dev_tables = nullptr;
    luzrt_material* dev_mats = nullptr;
    uint32_t* counters = nullptr;       // Counter array of the batch's traversal
    BatchTables tables{};
    BatchTlas* dev_tlas = nullptr;
    TlasJob* dev_jobs = nullptr;
    RefitJob* dev_refits = nullptr;
    uint32_t max_m = 1;
    std::vector<luzrt_blas*> refit_blas;
    std::vector<const void*> refit_verts;
    uint32_t refit_chunks = 0, refit_jobs = 0;
    bool refits_dirty = true;
    // rays
    uint64_t max_rays = 0;
    uint32_t total_cblk = 0, total_pblk = 0, total_ppx = 0;
    DevBuf hits;
    // ray-group culling (decided at creation): candidate masks (kinds per
    // group), live-row bits and counts, the live-row list
    bool cull = false;
    uint32_t cull_kinds = 0, groups_max = 0;
    uint64_t* cull_cand = nullptr;
    uint64_t* cull_bits = nullptr;
    uint32_t* cull_counts = nullptr;
    uint32_t* cull_rows = nullptr;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // launch configuration and workspace a captured graph was recorded with
    std::array<uint64_t, 7> graph_key{};
    hipEvent_t done = nullptr;
    bool pending = false;
    uint32_t frame = 0;
};

namespace {

void graph_release(luzrt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

luzrt_status wait_frame(luzrt_ctx* ctx, luzrt_batch* b) {
    if (b->pending) {
        LUZ_HIP(ctx, hipEventSynchronize(b->done));
        b->pending = false;
    }
    return LUZRT_OK;
}

// Scene table entries and the TLAS jobs/pointers for the current counts.
luzrt_status upload_tables(luzrt_ctx* ctx, luzrt_batch* b) {
    const uint32_t n = b->n;
    std::vector<BatchTlas> bt(n);
    std::vector<TlasJob> jobs(n);
    b->max_m = 1;
    for (uint32_t s = 0; s < n; ++s) {
        luzrt_tlas* t = b->desc[s].tlas;
        t->wide = false;   // batched frames build and trace the binary form only
        const InstIn* din = (const InstIn*)(b->dev_stage + b->inst_off[s]);
        jobs[s] = tlas_job(t, din, b->inst_n[s]);
        b->max_m = std::max(b->max_m, jobs[s].m);
        bt[s].nodes = b->inst_n[s] ? t->nodes : nullptr;
        bt[s].recs = t->recs;
        bt[s].position = b->sc[s].g.position;
        bt[s].normal = b->sc[s].g.normal;
        bt[s].ds = b->sc[s].ds;
        bt[s].act_list = b->sc[s].act_list;
        bt[s].width = b->sc[s].g.width;
        bt[s].ninst = b->inst_n[s];
        bt[s].slot_box = tlas_slot_boxes(t);
    }
    LUZ_HIP(ctx, hipMemcpyAsync(b->dev_tlas, bt.data(), n * sizeof(BatchTlas), hipMemcpyHostToDevice,
                                ctx->stream));
    LUZ_HIP(ctx, hipMemcpyAsync(b->dev_jobs, jobs.data(), n * sizeof(TlasJob), hipMemcpyHostToDevice,
                                ctx->stream));
    // the scene table carries the TLAS pointers for the primary kernel
    for (uint32_t s = 0; s < n; ++s) {
        b->sc[s].tlas_nodes = bt[s].nodes;
        b->sc[s].recs = bt[s].recs;
        b->sc[s].aux = b->desc[s].tlas->aux;
    }
    LUZ_HIP(ctx, hipMemcpyAsync((void*)b->tables.sc, b->sc.data(), n * sizeof(BatchSceneDev),
                                hipMemcpyHostToDevice, ctx->stream));
    LUZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LUZRT_OK;
}

luzrt_status upload_refits(luzrt_ctx* ctx, luzrt_batch* b) {
    std::vector<RefitJob> jobs;
    uint32_t chunks = 0;
    for (size_t i = 0; i < b->refit_blas.size(); ++i) {
        luzrt_blas* bl = b->refit_blas[i];
        if (bl->tri_count < 2) continue;   // a single-triangle root is rebuilt in issue_frame
        RefitJob J = refit_job(bl, b->refit_verts[i], nullptr);
        J.chunk_offset = chunks;
        chunks += refit_chunks(bl->tri_count);
        jobs.push_back(J);
    }
    if (b->dev_refits) LUZ_HIP(ctx, hipFree(b->dev_refits));
    b->dev_refits = nullptr;
    b->refit_jobs = (uint32_t)jobs.size();
    b->refit_chunks = chunks;
    if (!jobs.empty()) {
        LUZ_HIP(ctx, hipMalloc((void**)&b->dev_refits, jobs.size() * sizeof(RefitJob)));
        LUZ_HIP(ctx, hipMemcpy(b->dev_refits, jobs.data(), jobs.size() * sizeof(RefitJob),
                               hipMemcpyHostToDevice));
    }
    b->refits_dirty = false;
    return LUZRT_OK;
}

// The frame's launches, in order, on ctx->stream (captured or issued directly).
luzrt_status issue_frame(luzrt_ctx* ctx, luzrt_batch* b) {
    hipStream_t st = ctx->stream;
    uint32_t* error = build_error_flag(ctx);
    if (!error) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    LUZ_HIP(ctx, hipMemcpyAsync(b->dev_stage, b->host_stage, b->stage_bytes, hipMemcpyHostToDevice, st));
    luzrt_status s;
    if (b->refit_jobs &&
        (s = launch_refit_batched(ctx, b->dev_refits, b->refit_jobs, b->refit_chunks)) !=
            LUZRT_OK)
        return s;
    // BLASes of fewer than 2 triangles are not refit jobs: their single root
    // is rebuilt from the moved vertices as luzrt_blas_refit does
    for (size_t i = 0; i < b->refit_blas.size(); ++i)
        if (b->refit_blas[i]->tri_count < 2 &&
            (s = blas_refit_device(ctx, b->refit_blas[i], b->refit_verts[i])) != LUZRT_OK)
            return s;
    if ((s = launch_tlas_batched(ctx, b->dev_jobs, b->n, b->max_m)) != LUZRT_OK) return s;
    k_batch_init<<<1, 64, 0, st>>>(b->counters);
    if (b->total_ppx) {
        uint32_t groups = (uint32_t)ctx->num_cus * kPrimaryGroupsPerCu;
        if (groups == 0) groups = 1;
        k_bprimary<<<groups, kBlock, 0, st>>>(b->tables, (uint32_t*)ctx->spill.ptr, b->counters,
                                              primary_chunk(b->total_ppx / 64, groups));
    }
    const uint32_t tw = ctx->tile_width;
    k_bcompact_count<<<b->total_cblk, kBlock, 0, st>>>(b->tables, tw);
    k_bcompact_scan<<<b->n, 1024, 0, st>>>(b->tables);
    k_bray_base<<<1, 64, 0, st>>>(b->tables, b->counters, b->cull ? 16u : 1u);
    k_bcompact_write<<<b->total_cblk, kBlock, 0, st>>>(b->tables, tw);
    LUZ_HIP(ctx, hipGetLastError());
    if (b->cull) {
        const uint32_t gblocks = cdiv(b->groups_max, kBlock / 64);
        k_bcull<<<gblocks, kBlock, 0, st>>>(b->tables, b->dev_tlas, b->groups_max, b->cull_kinds, b->cull_cand,
                                           b->cull_bits, b->cull_counts);
        launch_cull_scan(st, b->cull_counts, b->groups_max, b->counters);
        k_bcull_rows<<<gblocks, kBlock, 0, st>>>(b->tables, b->groups_max, b->cull_bits, b->cull_counts,
                                                b->cull_rows, (uint8_t*)b->hits.ptr);
        LUZ_HIP(ctx, hipGetLastError());
    }
    TraceLaunch L{};
    if (b->cull) {
        L.live_rows = b->cull_rows;
        L.cand = b->cull_cand;
        L.cand_kinds = b->cull_kinds;
        L.scene_group_base = b->tables.gbase;
    }
    L.hits = (uint8_t*)b->hits.ptr;
    L.counters = b->counters;
    L.wide = false;
    L.nscenes = b->n;
    L.scene_ray_base = b->tables.ray_base;
    L.scene_tlas = b->dev_tlas;
    L.scene_active = b->tables.active;
    if ((s = launch_trace(ctx, L)) != LUZRT_OK) return s;
    k_bresolve<<<b->total_pblk, kBlock, 0, st>>>(b->tables, (const uint8_t*)b->hits.ptr);
    k_bshade<<<b->total_pblk, kBlock, 0, st>>>(b->tables, (const float*)ctx->lut.ptr);
    LUZ_HIP(ctx, hipGetLastError());
    return LUZRT_OK;
}

// Workspace that the frame's launches use, allocated before any capture.
luzrt_status prepare_frame(luzrt_ctx* ctx, luzrt_batch* b) {
    luzrt_status s;
    if ((s = ensure_lut(ctx)) != LUZRT_OK) return s;
    if (!build_error_flag(ctx)) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory");
    uint32_t groups = (uint32_t)ctx->num_cus * std::max<uint32_t>(ctx->groups_per_cu, kPrimaryGroupsPerCu);
    if (groups == 0) groups = 1;
    if ((s = ensure(ctx, ctx->spill, (size_t)groups * kBlock * kSpill * 4)) != LUZRT_OK) return s;
    // a culled frame's scenes start at multiples of 16 rays (k_bray_base)
    if ((s = ensure(ctx, b->hits, b->max_rays + 16 * ((size_t)b->n + 1))) != LUZRT_OK) return s;
    return LUZRT_OK;
}

// Compaction block counts of every scene depend on the tile width.
luzrt_status layout_blocks(luzrt_ctx* ctx, luzrt_batch* b) {
    const uint32_t n = b->n;
    const uint32_t tw = ctx->tile_width, th = 64u / tw;
    std::vector<BatchSceneDev>& sc = b->sc;
    std::vector<uint32_t> cblk(n + 1);
    uint64_t c = 0;
    for (uint32_t s = 0; s < n; ++s) {
        const uint32_t W = sc[s].g.width, H = sc[s].g.height;
        sc[s].ntiles_px = ((W + tw - 1) / tw) * ((H + th - 1) / th) * 64;
        sc[s].nblocks = cdiv(sc[s].ntiles_px, kBlock);
        cblk[s] = (uint32_t)c;
        c += sc[s].nblocks;
    }
    cblk[n] = (uint32_t)c;
    b->total_cblk = (uint32_t)c;
    LUZ_HIP(ctx, hipMemcpy((void*)b->tables.sc, sc.data(), n * sizeof(BatchSceneDev),
                           hipMemcpyHostToDevice));
    LUZ_HIP(ctx, hipMemcpy((void*)b->tables.cblk, cblk.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    return LUZRT_OK;
}

void batch_free(luzrt_batch* b) {
    graph_release(b);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->host_stage) (void)hipHostFree(b->host_stage);
    for (void* p : {(void*)b->dev_stage, (void*)b->dev_work, (void*)b->dev_tables, (void*)b->dev_mats,
                    (void*)b->dev_tlas, (void*)b->dev_jobs, (void*)b->dev_refits,
                    b->hits.ptr, (void*)b->cull_cand, (void*)b->cull_bits, (void*)b->cull_counts,
                    (void*)b->cull_rows})
        if (p) (void)hipFree(p);
    delete b;
}

}  // namespace

extern "C" {

luzrt_status luzrt_batch_create(luzrt_ctx* ctx, const luzrt_batch_scene* scenes, uint32_t count,
                                luzrt_batch** out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!out || (count && !scenes)) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_create: null argument");
    *out = nullptr;
    if (count == 0 || count > kMaxBatch)
        return fail(ctx, LUZRT_E_LIMIT, "batch of %u scenes outside 1..%u", count, kMaxBatch);
    if (ctx->precision != LUZRT_PRECISION_PLAIN)
        return fail(ctx, LUZRT_E_UNSUPPORTED, "batched frames use the plain precision form");
    LUZ_HIP(ctx, hipSetDevice(ctx->device));
    auto* b = new luzrt_batch();
    b->n = count;
    b->desc.assign(scenes, scenes + count);
    std::vector<DevScene> ds(count);
    uint64_t rays = 0;
    size_t work_bytes = 0, inst_bytes = 0;
    for (uint32_t s = 0; s < count; ++s) {
        const luzrt_batch_scene& d = scenes[s];
        luzrt_status st;
        auto bad = [&](luzrt_status code, const char* what) {
            batch_free(b);
            return fail(ctx, code, "batch scene %u: %s", s, what);
        };
        if (!d.tlas || !d.rgba32f_out || !d.ao_counts || (d.scene.num_lights > 0 && !d.shadow_counts))
            return bad(LUZRT_E_INVALID_ARG, "null TLAS or output");
        if ((st = fill_scene(ctx, &d.scene, &ds[s], true)) != LUZRT_OK) { batch_free(b); return st; }
        luzrt_gbuffer g;
        if ((st = check_gbuffer(ctx, &d.gbuffer, &g)) != LUZRT_OK) { batch_free(b); return st; }
        // the batched kernels (and the primary rays that write the planes)
        // address tightly packed planes with a position plane
        if (!gbuffer_packed(g))
            return bad(LUZRT_E_UNSUPPORTED, "batched frames take tightly packed G-buffer planes with a position plane");
        if (d.material_count && !d.materials) return bad(LUZRT_E_INVALID_ARG, "null material table");
        const uint64_t npix = (uint64_t)g.width * g.height;
        rays += npix * ds[s].r_total;
        work_bytes += (size_t)npix * 8 + ((size_t)cdiv(npix + 63 * g.width + 63 * g.height + 4096, kBlock) + 2) * 4;
        inst_bytes += ((size_t)d.tlas->capacity * sizeof(InstIn) + 255) & ~(size_t)255;
        b->cams.push_back(d.camera ? *d.camera : luzrt_camera{});
        b->has_cam.push_back(d.camera ? 1u : 0u);
        b->mats.emplace_back(d.materials, d.materials + d.material_count);
    }
    if (rays + 16ull * (count + 1) > 0xFFFFFFFFull) {
        batch_free(b);
        return fail(ctx, LUZRT_E_LIMIT, "%llu rays per frame exceed the 32-bit ray index",
                    (unsigned long long)rays);
    }
    // ray-group culling when every scene allows it (<= 128 TLAS slots, <= 64 rays
    // per pixel): the scenes' rays then go in 64-pixel groups
    uint64_t groups = 0;
    uint32_t kinds = 1, rmax = 1;
    bool cull = ctx->group_cull;
    for (uint32_t s = 0; s < count; ++s) {
        const uint64_t npix = (uint64_t)scenes[s].gbuffer.width * scenes[s].gbuffer.height;
        groups += (npix + kCullGroup - 1) / kCullGroup;
        kinds = std::max<uint32_t>(kinds, (uint32_t)scenes[s].scene.num_lights + 1);
        rmax = std::max<uint32_t>(rmax, ds[s].r_total);
        if (scenes[s].tlas->capacity > kMaxCullInst || ds[s].r_total > kMaxCullSlots) cull = false;
    }
    if (groups * kCullGroup > 0xFFFFFFFFull) cull = false;
    b->cull = cull;
    b->cull_kinds = kinds;
    b->groups_max = (uint32_t)groups;
    for (uint32_t s = 0; s < count; ++s)
        ds[s].ray_group = (cull || ctx->ray_order == LUZRT_RAY_ORDER_GROUPED) ? kRayGroup : 0;
    // two scenes must not share a TLAS (one build job per TLAS and frame), an
    // output, or a G-buffer plane that either of them writes (camera set)
    for (uint32_t s = 0; s < count; ++s)
        for (uint32_t t = 0; t < s; ++t) {
            const luzrt_batch_scene &a = scenes[s], &c = scenes[t];
            const char* what = nullptr;
            if (a.tlas == c.tlas) what = "the same TLAS";
            else if (a.rgba32f_out == c.rgba32f_out || a.ao_counts == c.ao_counts ||
                     (a.shadow_counts && a.shadow_counts == c.shadow_counts))
                what = "an output buffer";
            else if ((a.camera || c.camera) &&
                     (a.gbuffer.position == c.gbuffer.position || a.gbuffer.normal == c.gbuffer.normal ||
                      a.gbuffer.albedo == c.gbuffer.albedo || a.gbuffer.material == c.gbuffer.material ||
                      a.gbuffer.emission == c.gbuffer.emission))
                what = "a G-buffer plane that a camera writes";
            if (what) {
                batch_free(b);
                return fail(ctx, LUZRT_E_INVALID_ARG, "batch scenes %u and %u share %s", t, s, what);
            }
        }
    b->max_rays = rays;
    const size_t scene_bytes = ((size_t)count * sizeof(DevScene) + 255) & ~(size_t)255;
    b->stage_bytes = scene_bytes + inst_bytes;
    hipError_t e = hipHostMalloc((void**)&b->host_stage, b->stage_bytes, hipHostMallocDefault);
    if (e != hipSuccess) { batch_free(b); return hip_fail(ctx, e, "hipHostMalloc(batch staging)"); }
    std::memset(b->host_stage, 0, b->stage_bytes);
    std::memcpy(b->host_stage, ds.data(), count * sizeof(DevScene));
    size_t off = scene_bytes;
    for (uint32_t s = 0; s < count; ++s) {
        b->inst_off.push_back(off);
        off += ((size_t)scenes[s].tlas->capacity * sizeof(InstIn) + 255) & ~(size_t)255;
    }
    b->inst_n.assign(count, 0);
    b->graph_n.assign(count, ~0u);
    b->staged.assign(count, 0);
    size_t nmats = 0;
    for (auto& m : b->mats) nmats += m.size();
    const size_t tables_bytes = count * sizeof(BatchSceneDev) + (count + 1) * 4 * 3 + count * 4 +
                                (count + 1) * 4 * 2 + 256;
#define B_ALLOC(ptr, bytes)                                                             \
    do {                                                                                \
        hipError_t ae = hipMalloc((void**)&(ptr), (bytes) ? (bytes) : 256);             \
        if (ae != hipSuccess) { batch_free(b); return hip_fail(ctx, ae, "hipMalloc(batch)"); } \
    } while (0)
    B_ALLOC(b->dev_stage, b->stage_bytes);
    B_ALLOC(b->dev_work, work_bytes);
    B_ALLOC(b->dev_tables, tables_bytes);
    B_ALLOC(b->dev_mats, nmats * sizeof(luzrt_material));
    B_ALLOC(b->counters, kCounterBytes);
    B_ALLOC(b->dev_tlas, count * sizeof(BatchTlas));
    B_ALLOC(b->dev_jobs, count * sizeof(TlasJob));
    if (b->cull) {
        B_ALLOC(b->cull_cand, groups * kinds * kCandWords * 8);
        B_ALLOC(b->cull_bits, groups * 8);
        B_ALLOC(b->cull_counts, ((groups + 3) & ~uint64_t(3)) * 4);
        B_ALLOC(b->cull_rows, groups * rmax * 4);
    }
#undef B_ALLOC
    // the error latch in the counters starts clear; frames never clear it
    e = hipMemset(b->counters, 0, kCounterBytes);
    if (e != hipSuccess) { batch_free(b); return hip_fail(ctx, e, "hipMemset(batch counters)"); }
    e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e != hipSuccess) { batch_free(b); return hip_fail(ctx, e, "hipEventCreate"); }

    // tables
    uint8_t* tp = b->dev_tables;
    BatchSceneDev* dsc = (BatchSceneDev*)tp;
    tp += count * sizeof(BatchSceneDev);
    b->tables.sc = dsc;
    b->tables.cblk = (const uint32_t*)tp; tp += (count + 1) * 4;
    b->tables.pblk = (const uint32_t*)tp; tp += (count + 1) * 4;
    b->tables.ppx = (const uint32_t*)tp; tp += (count + 1) * 4;
    b->tables.active = (uint32_t*)tp; tp += count * 4;
    b->tables.ray_base = (uint32_t*)tp; tp += (count + 1) * 4;
    b->tables.gbase = (uint32_t*)tp;
    b->tables.n = count;
    b->sc.resize(count);
    std::vector<BatchSceneDev>& sc = b->sc;
    std::vector<uint32_t> pblk(count + 1), ppx(count + 1);
    uint8_t* wp = b->dev_work;
    luzrt_material* mp = b->dev_mats;
    uint64_t pb = 0, px = 0;
    for (uint32_t s = 0; s < count; ++s) {
        const luzrt_batch_scene& d = scenes[s];
        BatchSceneDev& S = sc[s];
        std::memset(&S, 0, sizeof(S));
        (void)check_gbuffer(ctx, &d.gbuffer, &S.g);   // validated above: the layout filled in
        S.gout.width = d.gbuffer.width;
        S.gout.height = d.gbuffer.height;
        S.gout.position = const_cast<float*>(d.gbuffer.position);
        S.gout.normal = const_cast<float*>(d.gbuffer.normal);
        S.gout.albedo = const_cast<uint32_t*>(d.gbuffer.albedo);
        S.gout.material = const_cast<uint32_t*>(d.gbuffer.material);
        S.gout.emission = const_cast<uint32_t*>(d.gbuffer.emission);
        S.cam = b->cams[s];
        S.has_camera = b->has_cam[s];
        S.ds = (const DevScene*)(b->dev_stage + (size_t)s * sizeof(DevScene));
        S.mats = mp;
        S.nmats = (uint32_t)b->mats[s].size();
        if (S.nmats)
            LUZ_HIP(ctx, hipMemcpy(mp, b->mats[s].data(), S.nmats * sizeof(luzrt_material),
                                   hipMemcpyHostToDevice));
        mp += S.nmats;
        S.shadow = d.shadow_counts;
        S.ao = d.ao_counts;
        S.out = (float4*)d.rgba32f_out;
        const uint32_t W = d.gbuffer.width, H = d.gbuffer.height;
        S.npix = W * H;
        S.act_list = (uint32_t*)wp; wp += (size_t)S.npix * 4;
        S.act_index = (uint32_t*)wp; wp += (size_t)S.npix * 4;
        S.block_counts = (uint32_t*)wp;
        // room for the block counts of any tile width (8..64 wide tiles pad <= 63 px per side)
        wp += ((size_t)cdiv((uint64_t)S.npix + 63 * W + 63 * H + 4096, kBlock) + 2) * 4;
        pblk[s] = (uint32_t)pb;
        pb += cdiv(S.npix, kBlock);
        ppx[s] = (uint32_t)px;
        if (S.has_camera) px += (uint64_t)((W + 7) / 8) * ((H + 7) / 8) * 64;
    }
    pblk[count] = (uint32_t)pb;
    ppx[count] = (uint32_t)px;
    b->total_pblk = (uint32_t)pb;
    b->total_ppx = (uint32_t)px;
    LUZ_HIP(ctx, hipMemcpy(dsc, sc.data(), count * sizeof(BatchSceneDev), hipMemcpyHostToDevice));
    LUZ_HIP(ctx, hipMemcpy((void*)b->tables.pblk, pblk.data(), (count + 1) * 4, hipMemcpyHostToDevice));
    LUZ_HIP(ctx, hipMemcpy((void*)b->tables.ppx, ppx.data(), (count + 1) * 4, hipMemcpyHostToDevice));
    luzrt_status st = layout_blocks(ctx, b);
    if (st != LUZRT_OK) { batch_free(b); return st; }
    *out = b;
    return LUZRT_OK;
}

luzrt_status luzrt_batch_add_refit(luzrt_ctx* ctx, luzrt_batch* b, luzrt_blas* blas,
                                   const void* vertices) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b || !blas || !vertices) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_add_refit: null argument");
    if (std::find(b->refit_blas.begin(), b->refit_blas.end(), blas) != b->refit_blas.end())
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_add_refit: BLAS registered twice");
    luzrt_status s = wait_frame(ctx, b);
    if (s != LUZRT_OK) return s;
    b->refit_blas.push_back(blas);
    b->refit_verts.push_back(vertices);
    b->refits_dirty = true;
    graph_release(b);
    return LUZRT_OK;
}

luzrt_status luzrt_batch_set_instances(luzrt_ctx* ctx, luzrt_batch* b, uint32_t scene,
                                       const luzrt_instance* instances, uint32_t count) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b || scene >= b->n) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_set_instances: bad batch or scene");
    // the staging buffer is read by the previous frame's copy
    luzrt_status s = wait_frame(ctx, b);
    if (s != LUZRT_OK) return s;
    std::vector<InstIn> host;
    if ((s = pack_instances(ctx, b->desc[scene].tlas, instances, count, host)) != LUZRT_OK) return s;
    if (!host.empty())
        std::memcpy(b->host_stage + b->inst_off[scene], host.data(), host.size() * sizeof(InstIn));
    b->inst_n[scene] = (uint32_t)host.size();
    b->staged[scene] = 1;
    return LUZRT_OK;
}

luzrt_status luzrt_batch_frame(luzrt_ctx* ctx, luzrt_batch* b, uint32_t frame, int use_graph) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_frame: null batch");
    if (ctx->precision != LUZRT_PRECISION_PLAIN)
        return fail(ctx, LUZRT_E_UNSUPPORTED, "batched frames use the plain precision form");
    for (uint32_t s = 0; s < b->n; ++s)
        if (!b->staged[s])
            return fail(ctx, LUZRT_E_INVALID_ARG, "batch scene %u has no staged instances", s);
    luzrt_status st = wait_frame(ctx, b);
    if (st != LUZRT_OK) return st;
    LUZ_HIP(ctx, hipSetDevice(ctx->device));
    // the frame index of every scene's sample sequence
    for (uint32_t s = 0; s < b->n; ++s)
        ((DevScene*)b->host_stage)[s].s.frame = (int32_t)(b->desc[s].scene.frame + frame);
    for (luzrt_blas* bl : b->refit_blas) bl->wide_valid = false;
    // a graph holds the launch configuration and the context's traversal
    // spill buffer by value: a single-scene call may reallocate that buffer,
    // so its address is part of the key
    auto key = [&]() {
        return std::array<uint64_t, 7>{ctx->groups_per_cu, ctx->refill_threshold, ctx->stack_entries,
                                       ctx->tile_width, (uint64_t)(uintptr_t)ctx->spill.ptr,
                                       (uint64_t)ctx->spill.bytes, ctx->max_steps};
    };
    bool same = b->exec && key() == b->graph_key && b->inst_n == b->graph_n;
    if (!same || !use_graph) {
        graph_release(b);
        if (ctx->tile_width != b->graph_key[3] || !b->graph_key[3])
            if ((st = layout_blocks(ctx, b)) != LUZRT_OK) return st;
        if (b->refits_dirty && (st = upload_refits(ctx, b)) != LUZRT_OK) return st;
        if ((st = upload_tables(ctx, b)) != LUZRT_OK) return st;
        if ((st = prepare_frame(ctx, b)) != LUZRT_OK) return st;
        b->graph_key = key();
        b->graph_n = b->inst_n;
    }
    ctx->release_clean = false;   // the frame refits BLASes and reads TLASes
    if (use_graph) {
        if (!b->exec) {
            LUZ_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            st = issue_frame(ctx, b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (st != LUZRT_OK) { if (g) (void)hipGraphDestroy(g); return st; }
            if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
            b->graph = g;
            LUZ_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        LUZ_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    } else {
        if ((st = issue_frame(ctx, b)) != LUZRT_OK) return st;
    }
    LUZ_HIP(ctx, hipEventRecord(b->done, ctx->stream));
    b->pending = true;
    b->frame = frame;
    return LUZRT_OK;
}

luzrt_status luzrt_batch_rays(luzrt_ctx* ctx, luzrt_batch* b, uint64_t* total_rays) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b || !total_rays) return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_batch_rays: null argument");
    luzrt_status s = wait_frame(ctx, b);
    if (s != LUZRT_OK) return s;
    uint32_t c[kCntCount];
    LUZ_HIP(ctx, hipMemcpy(c, b->counters, sizeof(c), hipMemcpyDeviceToHost));
    *total_rays = c[kCntRays];
    uint32_t build = 0;
    uint32_t* flag = build_error_flag(ctx);
    if (!flag) return fail(ctx, LUZRT_E_OUT_OF_MEMORY, "no device memory for the error flag");
    LUZ_HIP(ctx, hipMemcpy(&build, flag, 4, hipMemcpyDeviceToHost));
    if (build) {
        LUZ_HIP(ctx, hipMemset(flag, 0, 4));
        return fail(ctx, LUZRT_E_DEVICE, "batched BVH build or refit reported error %u", build);
    }
    if (c[kCntError]) {
        LUZ_HIP(ctx, hipMemset(b->counters + kCntError, 0, 4));
        return fail(ctx, LUZRT_E_LIMIT,
                    "a ray exceeded the traversal stack or step limit in a frame since the last check "
                    "(such rays are written as misses)");
    }
    return LUZRT_OK;
}

luzrt_status luzrt_batch_destroy(luzrt_ctx* ctx, luzrt_batch* b) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!b) return LUZRT_OK;
    (void)wait_frame(ctx, b);
    batch_free(b);
    return LUZRT_OK;
}

}  // extern "C"
