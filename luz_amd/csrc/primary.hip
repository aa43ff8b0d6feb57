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
#include "traverse.hpp"

namespace luz {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_primary(const Node* tlas_nodes, const InstRec* recs,
                                                    const InstAux* aux, luzrt_camera cam,
                                                    const luzrt_material* mats, uint32_t nmats,
                                                    luzrt_gbuffer_out out, uint32_t* spill,
                                                    uint32_t* error_flag, bool empty) {
    __shared__ uint32_t lds[kShort * kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = gridDim.x * kBlock;
    Stack st;
    st.ls = lds + tid;
    st.lds_stride = kBlock;
    st.gs = spill + (size_t)blockIdx.x * kBlock + tid;
    st.g_stride = total;
    const uint32_t W = out.width, H = out.height;
    const uint32_t tiles_x = (W + 7) / 8;
    const uint32_t ntiles_px = tiles_x * ((H + 7) / 8) * 64;
    const V3 eye = v3(cam.eye[0], cam.eye[1], cam.eye[2]);
    for (uint32_t g = blockIdx.x * kBlock + tid; g < ntiles_px; g += total) {
        uint32_t tile = g >> 6, t = g & 63u;
        uint32_t x = (tile % tiles_x) * 8 + (t & 7u), y = (tile / tiles_x) * 8 + (t >> 3);
        bool inside = x < W && y < H;
        Best best;
        best.found = false;
        best.t = cam.tmax;
        V3 d = v3(0.0f, 0.0f, 0.0f);
        if (inside && !empty) {
            float nx = ((float)x + 0.5f) / (float)W * 2.0f - 1.0f;
            float ny = ((float)y + 0.5f) / (float)H * 2.0f - 1.0f;
            d = v3(cam.forward[0] + nx * cam.right[0] + ny * cam.up[0],
                   cam.forward[1] + nx * cam.right[1] + ny * cam.up[1],
                   cam.forward[2] + nx * cam.right[2] + ny * cam.up[2]);
            RayState r;
            ray_begin(r, st, eye, d, cam.tmin, cam.tmax);
            for (;;) {
                int res = ray_step<true>(r, st, tlas_nodes, recs, aux, &best);
                if (res == kStepDone) break;
                if (res == kStepOverflow) { atomicOr(error_flag, 1u); break; }
            }
        }
        if (!inside) continue;
        size_t p = (size_t)y * W + x;
        if (!best.found) {
            for (int k = 0; k < 3; ++k) { out.position[p * 3 + k] = 0.0f; out.normal[p * 3 + k] = 0.0f; }
            out.albedo[p] = 0; out.material[p] = 0; out.emission[p] = 0;
            continue;
        }
        const InstAux& ax = aux[best.inst_slot];
        V3 P = vadd(eye, vmul(d, best.t));
        float bA = __fdiv_rn(best.U, best.det), bB = __fdiv_rn(best.V, best.det),
              bC = __fdiv_rn(best.W, best.det);
        const float* nn = ax.nrms + (size_t)best.tri_slot * 9;
        V3 no = v3(nn[0] * bA + nn[3] * bB + nn[6] * bC,
                   nn[1] * bA + nn[4] * bB + nn[7] * bC,
                   nn[2] * bA + nn[5] * bB + nn[8] * bC);
        const float* m = ax.nmat;
        V3 nw = v3(m[0] * no.x + m[1] * no.y + m[2] * no.z,
                   m[3] * no.x + m[4] * no.y + m[5] * no.z,
                   m[6] * no.x + m[7] * no.y + m[8] * no.z);
        float ln = vlen(nw);
        V3 N = ln == 0.0f ? v3(0.0f, 1.0f, 0.0f)
                          : v3(__fdiv_rn(nw.x, ln), __fdiv_rn(nw.y, ln), __fdiv_rn(nw.z, ln));
        out.position[p * 3 + 0] = P.x; out.position[p * 3 + 1] = P.y; out.position[p * 3 + 2] = P.z;
        out.normal[p * 3 + 0] = N.x; out.normal[p * 3 + 1] = N.y; out.normal[p * 3 + 2] = N.z;
        if (ax.custom_index < nmats) {
            luzrt_material mt = mats[ax.custom_index];
            out.albedo[p] = mt.albedo; out.material[p] = mt.material; out.emission[p] = mt.emission;
        } else {
            out.albedo[p] = 0xFFFFFFFFu;
            out.material[p] = 0xFFFF0080u;  // roughness 0.5, metallic 0, occlusion 1
            out.emission[p] = 0xFF000000u;
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
    luzrt_status s = ensure(ctx, ctx->counters, 64);
    if (s != LUZRT_OK) return s;
    // the material table lives in device memory for the duration of the call
    size_t mbytes = (size_t)(nmats ? nmats : 1) * sizeof(luzrt_material);
    static_assert(sizeof(luzrt_material) == 16, "material layout");
    luzrt_material* dmats = nullptr;
    LUZ_HIP(ctx, hipMallocAsync((void**)&dmats, mbytes, ctx->stream));
    if (nmats)
        LUZ_HIP(ctx, hipMemcpyAsync(dmats, materials, (size_t)nmats * sizeof(luzrt_material),
                                    hipMemcpyHostToDevice, ctx->stream));
    uint32_t groups = (uint32_t)ctx->num_cus * 4;
    if (groups == 0) groups = 1;
    if ((s = ensure(ctx, ctx->spill, (size_t)groups * 256 * kSpill * 4)) != LUZRT_OK) return s;
    uint32_t* counters = (uint32_t*)ctx->counters.ptr;
    LUZ_HIP(ctx, hipMemsetAsync(counters + kCntError, 0, 4, ctx->stream));
    stage_begin(ctx, kStagePrimary);
    k_primary<<<groups, 256, 0, ctx->stream>>>(tlas->nodes, tlas->recs, tlas->aux, *cam, dmats, nmats,
                                              *out, (uint32_t*)ctx->spill.ptr, counters + kCntError,
                                              tlas->count == 0);
    stage_end(ctx, kStagePrimary);
    LUZ_HIP(ctx, hipGetLastError());
    LUZ_HIP(ctx, hipFreeAsync(dmats, ctx->stream));
    return LUZRT_OK;
}
