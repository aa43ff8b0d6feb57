// primary.hip — primary-ray G-buffer producer on gfx950.
//
// The reference rasterises the G-buffer (opaque.vert:21-31, opaque.frag:21-58,
// DeferredRenderer.cpp:254-262); the lighting path's inputs are defined here
// "by primary rays": a persistent closest-hit traversal over the same TLAS,
// writing world position, normalised interpolated vertex normal (through the
// instance's inverse-transpose), and the material of the hit instance.
// Ties in t are broken by (instance input index, primitive id), so the result
// does not depend on traversal order.
#include "internal.hpp"
#include "lightpath.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

// Clears the tile counter of the call with a kernel (a caller may capture
// the call in a graph, where memset nodes are avoided).  The error latch
// kCntError is left set until the host reports it.
__global__ void k_primary_init(uint32_t* counters) {
    if (threadIdx.x == 0) counters[kCntPrimary] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_primary(const Node* tlas_nodes, const InstRec* recs,
                                                    const InstAux* aux, luzrt_camera cam,
                                                    const luzrt_material* mats, uint32_t nmats,
                                                    luzrt_gbuffer_out out, uint32_t* spill,
                                                    uint32_t* counters, uint32_t chunk, bool empty) {
    __shared__ uint32_t lds[kShort * kBlock];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t total = gridDim.x * kBlock;
    Stack st;
    st.init(lds + tid, kBlock, spill + (size_t)blockIdx.x * kBlock + tid, total);
    const uint32_t W = out.width, H = out.height;
    const uint32_t tiles_x = (W + 7) / 8;
    const uint32_t ntiles = tiles_x * ((H + 7) / 8);
    // each wave takes the next `chunk` 8x8 tiles (lane = pixel of a tile) until
    // none is left: waves that drew cheap tiles take more of them
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&counters[kCntPrimary], chunk);
        base = __shfl(base, 0);
        if (base >= ntiles) break;
        const uint32_t end = min(base + chunk, ntiles);
        for (uint32_t tile = base; tile < end; ++tile) {
            uint32_t x = (tile % tiles_x) * 8 + (lane & 7u), y = (tile / tiles_x) * 8 + (lane >> 3);
            if (x < W && y < H)
                primary_pixel(tlas_nodes, recs, aux, cam, mats, nmats, out, x, y, st,
                              &counters[kCntError], empty);
        }
    }
}

}  // namespace
}  // namespace luz

using namespace luz;

extern "C" luzrt_status luzrt_primary_gbuffer(luzrt_ctx* ctx, const luzrt_tlas* tlas,
                                              const luzrt_camera* cam,
                                              const luzrt_material* materials, uint32_t nmats,
                                              const luzrt_gbuffer_out* out) {
    if (!ctx) return LUZRT_E_INVALID_ARG;
    if (!tlas || !cam || !out || (nmats && !materials) || !out->position || !out->normal ||
        !out->albedo || !out->material || !out->emission)
        return fail(ctx, LUZRT_E_INVALID_ARG, "luzrt_primary_gbuffer: null argument");
    if (out->width == 0 || out->height == 0 || out->width > 16384 || out->height > 16384)
        return fail(ctx, LUZRT_E_LIMIT, "luzrt_primary_gbuffer: size %ux%u outside 1..16384",
                    out->width, out->height);
    luzrt_status s = ensure(ctx, ctx->counters, kCounterBytes);
    if (s != LUZRT_OK) return s;
    if ((s = tlas_acquire(ctx, tlas)) != LUZRT_OK) return s;
    // the material table lives in device memory for the duration of the call
    size_t mbytes = (size_t)(nmats ? nmats : 1) * sizeof(luzrt_material);
    static_assert(sizeof(luzrt_material) == 16, "material layout");
    luzrt_material* dmats = nullptr;
    LUZ_HIP(ctx, hipMallocAsync((void**)&dmats, mbytes, ctx->stream));
    if (nmats)
        LUZ_HIP(ctx, hipMemcpyAsync(dmats, materials, (size_t)nmats * sizeof(luzrt_material),
                                    hipMemcpyHostToDevice, ctx->stream));
    uint32_t groups = (uint32_t)ctx->num_cus * kPrimaryGroupsPerCu;
    if (groups == 0) groups = 1;
    if ((s = ensure(ctx, ctx->spill, (size_t)groups * 256 * kSpill * 4)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    stage_begin(ctx, kStagePrimary);
    k_primary_init<<<1, 64, 0, ctx->stream>>>(counters);
    const uint32_t ntiles = ((out->width + 7) / 8) * ((out->height + 7) / 8);
    k_primary<<<groups, 256, 0, ctx->stream>>>(tlas->nodes, tlas->recs, tlas->aux, *cam, dmats, nmats,
                                              *out, (uint32_t*)ctx->spill.ptr, counters,
                                              primary_chunk(ntiles, groups), tlas->count == 0);
    stage_end(ctx, kStagePrimary);
    LUZ_HIP(ctx, hipGetLastError());
    LUZ_HIP(ctx, hipFreeAsync(dmats, ctx->stream));
    return LUZRT_OK;
}
