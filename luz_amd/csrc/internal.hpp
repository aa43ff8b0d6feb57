// internal.hpp — host-side objects behind the luzrt C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/luzrt.h"
#include "device_math.hpp"

namespace luz {

// A grow-only device allocation reused across calls.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum Stage { kStageBuild, kStageCompact, kStageRaygen, kStageTrace, kStageResolve, kStageShade,
             kStagePrimary, kStageDenoise, kStagePost, kStageCount };

}  // namespace luz

struct luzrt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // TLAS builds run on a stream of their own (default priority), after all
    // earlier work of `stream` (build_mark); the first later call that reads
    // the TLAS waits for its build (tlas_acquire), so the build overlaps the
    // work between the two on `stream`.
    hipStream_t build_stream = nullptr;
    hipEvent_t build_mark = nullptr;
    int pending_builds = 0;        // TLASes whose build no call has waited for
    // Recorded on `stream` after a visibility call's traversal; while
    // release_clean holds, nothing enqueued on `stream` since reads a TLAS or
    // writes a BLAS, so the next TLAS build waits for this event instead of
    // all of `stream` (it then overlaps the call's resolve or shading).
    hipEvent_t tlas_release = nullptr;
    bool release_clean = false;
    std::string error;
    int num_cus = 0;
    // workspace
    luz::DevBuf hits, active, act_index, act_rec, counters, spill, lut, scene, vis_shadow,
        vis_ao;
    // ray-group culling of visibility calls: candidate masks, live-row masks
    // and counts, the live-row list
    luz::DevBuf cull_cand, cull_bits, cull_rows;
    bool group_cull = true;        // luzrt_set_group_culling
    bool last_culled = false;      // the last visibility call traced live rows only
    // timing
    hipEvent_t ev[luz::kStageCount][2] = {};
    bool ev_used[luz::kStageCount] = {};
    bool timing = true;            // luzrt_set_stage_timing
    uint64_t last_rays = 0, last_active = 0;
    // traversal launch configuration
    uint32_t groups_per_cu = 6;      // launch shape chosen by the autotuner on MI355X
    uint32_t refill_threshold = 48;
    uint32_t stack_entries = 16;   // LDS stack entries per lane: 8, 16 or 32
    uint32_t tile_width = 8;       // pixel tiles of 64: 8x8, 16x4 or 32x2
    uint32_t ray_order = 0;        // LUZRT_RAY_ORDER_*: automatic, slot-major or 64-pixel groups
    uint32_t max_steps = 1u << 24; // traversal steps per ray before it is abandoned (kMaxSteps)
    // LUZRT_PRECISION_* mode: plain binary FP32 path or the compact form
    int precision = 0;
    // device address of the file-scope scratch of luzrt_trace_visibility_static
    // on this context's device (trace.hip), found once at luzrt_create
    void* static_scratch = nullptr;
    hipError_t static_scratch_error = hipSuccess;
    // pinned staging ring of small host-to-device uploads (scene blocks, TLAS
    // instances): stage_upload
    struct Staging {
        void* host = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int kStagingSlots = 4;
    Staging staging[kStagingSlots];
    int staging_next = 0;
};

struct luzrt_blas {
    uint32_t tri_count = 0;
    uint32_t node_count = 0;
    uint32_t vertex_count = 0;
    uint32_t stride = 0;
    luz::Node* nodes = nullptr;
    float* tris = nullptr;          // 9 floats per slot
    float* nrms = nullptr;          // 9 floats per slot
    uint32_t* prims = nullptr;      // slot -> primitive
    uint32_t* slot_idx = nullptr;   // 3 vertex indices per slot (refit)
    uint32_t* leaf_parent = nullptr;
    uint32_t* counters = nullptr;   // refit arrival counters (node_count)
    uint32_t* node_end = nullptr;   // per internal node: the far end of its leaf range
    uint32_t* span = nullptr;       // refit hand-over list: [0] = count, then node ids (tri_count)
    int refs = 0;                   // live TLASes built over this BLAS
    // precision form: 8-wide quantised nodes, built on demand from the binary tree
    luz::Node8* wide = nullptr;
    uint32_t wide_count = 0;
    float* tris8 = nullptr;          // 9 floats per wide leaf
    uint32_t* prims8 = nullptr;
    bool wide_valid = false;
};

struct luzrt_tlas {
    uint32_t capacity = 0;
    uint32_t count = 0;            // instances in the structure (after mask/empty filtering)
    uint32_t node_count = 0;
    luz::Node* nodes = nullptr;
    luz::InstRec* recs = nullptr;  // per slot, binary BLASes
    luz::InstRec* recs_w = nullptr;  // per slot, 8-wide BLASes (compact form)
    luz::InstAux* aux = nullptr;   // per slot
    float* w2o = nullptr;          // per slot, 12 floats (inspection copy)
    uint32_t* order = nullptr;     // slot -> input index
    void* input = nullptr;         // uploaded instance inputs
    void* scratch = nullptr;
    std::vector<const luzrt_blas*> referenced;
    bool wide = false;             // records point to 8-wide BLASes (compact form)
    hipEvent_t built = nullptr;    // recorded on the build stream after each build
    bool build_pending = false;    // no call on the context's stream has waited for it yet
};

namespace luz {

luzrt_status fail(luzrt_ctx* ctx, luzrt_status s, const char* fmt, ...);
luzrt_status hip_fail(luzrt_ctx* ctx, hipError_t e, const char* what);
luzrt_status ensure(luzrt_ctx* ctx, DevBuf& b, size_t bytes);
// Zeroes `bytes` bytes at `p` on the context's stream with a kernel rather
// than a memset node (calls on this path may be captured into a graph).
luzrt_status zero_fill(luzrt_ctx* ctx, void* p, size_t bytes);
void stage_begin(luzrt_ctx* ctx, Stage s, hipStream_t st = nullptr);   // null: the context's stream
void stage_end(luzrt_ctx* ctx, Stage s, hipStream_t st = nullptr);

#define LUZ_HIP(ctx, call)                                          \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) return ::luz::hip_fail(ctx, e_, #call); \
    } while (0)

// Host-side packed instance input for the TLAS build kernel (128 bytes).
struct alignas(16) InstIn {
    float o2w[12];
    float w2o[12];
    float nmat[9];
    uint32_t custom_index;
    uint32_t mask;
    uint32_t input_index;
    const Node* nodes;      // binary BLAS
    const float* tris;
    const float* nrms;
    const uint32_t* prims;
    const Node8* wnodes;    // 8-wide BLAS (compact form) or null
    const float* wtris;
};

// Launchers (each defined next to its kernels).
luzrt_status blas_build_device(luzrt_ctx* ctx, luzrt_blas* b, const void* vertices,
                               const uint32_t* indices);
luzrt_status blas_refit_device(luzrt_ctx* ctx, luzrt_blas* b, const void* vertices);
luzrt_status blas_wide_device(luzrt_ctx* ctx, luzrt_blas* b);
luzrt_status tlas_build_device(luzrt_ctx* ctx, luzrt_tlas* t, const InstIn* dev_in, uint32_t n,
                               hipStream_t st);
// Before the first kernel of a call that reads `tlas` on the context's stream:
// waits there for the TLAS's pending build (on the host while the stream is
// being captured).
luzrt_status tlas_acquire(luzrt_ctx* ctx, const luzrt_tlas* tlas);
// Waits on the host for the build stream.
luzrt_status build_stream_sync(luzrt_ctx* ctx);
// After the last kernel of a call that reads a TLAS: records tlas_release.
luzrt_status tlas_release(luzrt_ctx* ctx);
uint32_t* build_error_flag(luzrt_ctx* ctx);

// Host packing of TLAS instances (FP64 inverses, mask/empty filtering); updates
// the TLAS's bookkeeping (count, referenced BLASes) for the build that follows.
luzrt_status pack_instances(luzrt_ctx* ctx, luzrt_tlas* t, const luzrt_instance* inst,
                            uint32_t count, std::vector<InstIn>& out);

// One TLAS build job, as consumed by the single and the batched build kernels.
struct TlasJob {
    const InstIn* in;
    uint32_t n, m;          // instances, power of two >= n
    float* wbox;            // instance world boxes (6 floats each, input order)
    float* cbox[2];         // cluster boxes, ping-pong between clustering rounds
    uint32_t* cref[2];      // cluster node references, likewise
    Node* nodes;
    InstRec* recs;
    InstRec* recs_w;        // null unless the compact form is built
    InstAux* aux;
    float* w2o_out;
    uint32_t* order_out;
    float* slot_box;        // padded world box of each slot (6 floats), for ray-group culling
};
TlasJob tlas_job(luzrt_tlas* t, const InstIn* dev_in, uint32_t n);
// The padded world boxes of a built TLAS's slots (6 floats each, device).
const float* tlas_slot_boxes(const luzrt_tlas* t);
// Scratch bytes of a TLAS with room for `capacity` instances, and the dynamic
// LDS bytes of its build kernel for m (a power of two >= the instance count).
size_t tlas_scratch_bytes(uint32_t capacity);
size_t tlas_lds_bytes(uint32_t m);
luzrt_status launch_tlas_batched(luzrt_ctx* ctx, const TlasJob* dev_jobs, uint32_t njobs,
                                 uint32_t max_m);

// One BLAS refit job of a batched refit (binary form only).
// One BLAS of a refit: the boxes of every node from the leaf boxes, either
// recomputed from moved vertices (refit) or given per slot (build).
struct RefitJob {
    Node* nodes;
    uint32_t* leaf_parent;
    uint32_t* counters;
    const uint32_t* node_end;
    uint32_t* span;
    const uint32_t* slot_idx;
    float* tris;
    const uint8_t* verts;     // refit: moved vertices
    const float* leaf_box;    // build: 6 floats per slot
    uint32_t stride;
    uint32_t n;               // triangles (>= 2)
    uint32_t chunk_offset;    // first leaf chunk of this job among all jobs of the launch
    uint32_t pad;
};
luzrt_status launch_refit_batched(luzrt_ctx* ctx, const RefitJob* dev_jobs, uint32_t njobs,
                                  uint32_t total_chunks);
RefitJob refit_job(luzrt_blas* b, const void* vertices, const float* leaf_box);
uint32_t refit_chunks(uint32_t tris);

// Scene block as uploaded for one visibility call: the caller's luzrt_scene
// plus the ray-slot layout.  Each active (non-background) pixel owns r_total
// ray slots: shadow slots of light l are slot_base[l] .. slot_base[l] + n_l - 1,
// AO slots follow at r_shadow.  The ray index of slot j of active pixel a is
// ray_row() (lightpath.hpp).
struct DevScene {
    luzrt_scene s;
    uint32_t slot_base[LUZRT_MAX_LIGHTS];
    uint32_t r_shadow;
    uint32_t r_total;
    uint32_t ray_group;   // pixels per ray group (ray_row); 0: one group of all active pixels
    uint32_t pad;
};

// Device counters of one visibility call (ctx->counters).
enum Counter { kCntFetch = 0, kCntActive = 1, kCntError = 2, kCntRays = 3, kCntBuildError = 4,
               kCntPrimary = 5, kCntLive = 6, kCntCount = 8 };
// A traversal launch claims its rays in chunks from kClaimSegments counters,
// counter c owning chunks c, c + 8, c + 16, ..., each on a 128-byte line of
// its own after the Counter array: one counter took every wave's
// device-scope atomic (about 90k per 1080p frame, 1.9M per 64-scene batched
// frame).  A counter array is kCounterBytes long.
constexpr uint32_t kClaimSegments = 8;
constexpr uint32_t kClaimBase = 256;      // word index of segment 0's counter
constexpr uint32_t kClaimStride = 32;     // words between segment counters
constexpr size_t kCounterBytes = ((kClaimBase + kClaimSegments * kClaimStride) * 4 + 255) & ~size_t(255);

// Zeroes the claim counters (kClaimSegments threads of the calling block).
__device__ inline void clear_claims(uint32_t* counters, uint32_t tid) {
    if (tid < kClaimSegments) counters[kClaimBase + tid * kClaimStride] = 0u;
}

struct TraceLaunch {
    const Node* tlas_nodes;
    const InstRec* recs;
    const float4* ray_o;   // origin.xyz, tmin   (ray i at ray_o[i * ray_stride])
    const float4* ray_d;   // direction.xyz, tmax
    uint32_t ray_stride;   // in float4 units: 1 for split arrays, 2 for interleaved rays
    uint8_t* hits;
    uint32_t* counters;    // device Counter array; kCntRays read, kCntFetch zeroed before
    uint32_t* spill;       // per-lane overflow stack
    uint32_t max_steps;    // traversal steps per ray before it is abandoned (the context's limit)
    bool wide;             // instance records point to 8-wide BLASes
    // Multi-scene launch (nscenes > 0): ray i belongs to the scene s with
    // scene_ray_base[s] <= i < scene_ray_base[s + 1], is generated from that
    // scene's G-buffer (its ray i - scene_ray_base[s], with
    // A = scene_active[s]) and traced against scene_tlas[s]; tlas_nodes,
    // recs and the ray records are then unused.
    uint32_t nscenes;
    uint32_t tlas_nodes_count, recs_count;   // single-scene TLAS size (staged in LDS when small)
    const uint32_t* scene_ray_base;   // device, nscenes + 1 entries
    const struct BatchTlas* scene_tlas;
    const uint32_t* scene_active;     // device, nscenes entries
    const uint32_t* scene_group_base; // device, nscenes + 1: first culling group of each scene
    // Visibility launch with rays generated in place (gen): the ray of slot j
    // of active pixel a (ray_row order, A = counters[kCntActive]) is computed
    // from the pixel's record act_rec[2a], act_rec[2a + 1] (pixel_record) and
    // the scene block; ray_o and ray_d are then unused.
    bool gen;
    const struct DevScene* ds;
    const float4* act_rec;
    // Ray-group culling (gen launches, grouped order): the launch traces only
    // the live rows live_rows[0 .. counters[kCntLive]) (a row: slot j of the 64
    // active pixels of group g, encoded g << 6 | j), and a ray of kind k (its
    // light, or num_lights for AO) starts from the candidate slots of mask
    // words cand[(g * cand_kinds + k) * kCandWords ..] instead of the TLAS root.
    // Null: no culling.
    const uint32_t* live_rows;
    const uint64_t* cand;
    const float* slot_box;   // the TLAS slots' padded world boxes (6 floats each)
    uint32_t cand_kinds;
};

// One scene of a multi-scene launch: its TLAS (nodes null: empty TLAS) and
// the G-buffer, scene block and active-pixel list its rays are generated from.
struct BatchTlas {
    const Node* nodes;
    const InstRec* recs;
    const float* position;
    const float* normal;
    const struct DevScene* ds;
    const uint32_t* act_list;
    uint32_t width;
    uint32_t ninst;          // TLAS slots of the frame (ray-group culling)
    const float* slot_box;   // their padded world boxes (6 floats each)
};
constexpr uint32_t kMaxBatch = LUZRT_MAX_BATCH;
constexpr uint32_t kRayGroup = 64;   // pixels per ray group of the grouped ray order (ray_row)
// Persistent workgroups per CU of the primary-ray kernels (92 VGPRs and 16 KB
// of LDS stacks per workgroup leave room for five waves per SIMD).  Their waves
// claim 8x8 pixel tiles from the kCntPrimary counter, `primary_chunk` tiles
// per claim: about 16 claims per wave, so that the claims of a large batch
// do not queue on the one counter.
constexpr uint32_t kPrimaryGroupsPerCu = 5;
inline uint32_t primary_chunk(uint64_t ntiles, uint32_t groups) {
    const uint64_t waves = (uint64_t)groups * 4;
    const uint64_t c = ntiles / (waves * 16);
    return c < 1 ? 1u : (c > 64 ? 64u : (uint32_t)c);
}
luzrt_status launch_trace(luzrt_ctx* ctx, const TraceLaunch& L);
// Exclusive scan in place of n live-row counts (16-byte aligned), the total to
// counters[kCntLive] (one workgroup, trace.hip).
void launch_cull_scan(hipStream_t st, uint32_t* counts, uint32_t n, uint32_t* counters);
// Finds the device address of the no-allocation entry point's scratch array
// on the context's device (current at the call); records a failure in the
// context for that entry point to report.
void static_scratch_init(luzrt_ctx* ctx);
// luzrt_trace_visibility's work: compaction, traversal with rays generated in
// place and, if `resolve`, the count planes `shadow` and `ao`.  Without it the
// hit bytes stay in ctx->hits (ray_row order over ctx->active, the active
// count in ctx->counters) for a pass that reads them directly.
luzrt_status visibility_core(luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_gbuffer* g,
                             const luzrt_scene* scene, uint8_t* shadow, uint8_t* ao, bool resolve);

// Validates a scene and fills its slot layout.  For a visibility call
// (`trace`) the lights of a scene that is not LUZRT_SHADOW_RAYTRACED own no
// shadow slots and their sample counts read 0 in the block: the reference
// packs numShadowSamples = 0 in those scenes (GPUScene.cpp:248) and casts
// shadow rays only in ray-traced ones (light.frag:141).
luzrt_status fill_scene(luzrt_ctx* ctx, const luzrt_scene* scene, DevScene* ds, bool trace);
// Copies `bytes` from host memory `src` to device memory `dst` on the
// context's stream through a pinned staging buffer: `src` may be reused as soon
// as the call returns, and the call does not wait for the GPU (a staging buffer
// is reused once its previous copy has finished).
luzrt_status stage_upload(luzrt_ctx* ctx, void* dst, const void* src, size_t bytes,
                          hipStream_t st = nullptr);   // null: the context's stream
// fill_scene, then uploads the block to ctx->scene with the given ray group.
luzrt_status upload_scene(luzrt_ctx* ctx, const luzrt_scene* scene, DevScene* host_copy,
                          uint32_t ray_group = 0, bool trace = false);
// Ray group of a visibility call on `tlas` under the context's ray order.
uint32_t ray_group_for(const luzrt_ctx* ctx, const luzrt_tlas* tlas, const luzrt_scene* scene = nullptr);
// Validates a G-buffer descriptor and writes it with its layout filled in
// (zero strides and pitches replaced by the packed defaults; the depth
// pointer cleared when a position plane is given): the form kernels take.
luzrt_status check_gbuffer(luzrt_ctx* ctx, const luzrt_gbuffer* g, luzrt_gbuffer* out);
// A checked descriptor of tightly packed planes with a position plane.
bool gbuffer_packed(const luzrt_gbuffer& g);
luzrt_status ensure_lut(luzrt_ctx* ctx);

}  // namespace luz
