// gpuscene.cpp — GPUScene: one BLAS per mesh of a host scene and the TLAS
// over its mesh nodes, through the lighting path's C ABI (GPUScene.cpp:50-52,
// 108-138, 354-365).  The host scene model itself (scene.cpp, ingest.cpp)
// does not touch the GPU.
#include <cstdint>
#include <vector>

#include "../../include/luzrt.h"
#include "../../include/luzscene.h"

#include "scene_model.hpp"

using namespace luz;

struct luzgpu {
    luzrt_ctx* ctx = nullptr;
    std::vector<void*> vbs, ibs;
    std::vector<luzrt_blas*> blas;
    luzrt_tlas* tlas = nullptr;
};

extern "C" {

luzgpu* luzgpu_create(luzrt_ctx* ctx, const luzscene* s) {
    if (!ctx || !s) return nullptr;
    auto* g = new luzgpu();
    g->ctx = ctx;
    // GPUScene::AddMesh (GPUScene.cpp:108-138): upload, one BLAS per distinct mesh.
    for (const MeshAsset& m : s->meshes) {
        void *vb = nullptr, *ib = nullptr;
        size_t vbytes = m.vertices.size() * sizeof(MeshVertex), ibytes = m.indices.size() * 4;
        if (luzrt_malloc(ctx, vbytes, &vb) != LUZRT_OK || luzrt_malloc(ctx, ibytes, &ib) != LUZRT_OK ||
            luzrt_upload(ctx, vb, m.vertices.data(), vbytes) != LUZRT_OK ||
            luzrt_upload(ctx, ib, m.indices.data(), ibytes) != LUZRT_OK) {
            if (vb) luzrt_free(ctx, vb);
            if (ib) luzrt_free(ctx, ib);
            luzgpu_destroy(g);
            return nullptr;
        }
        g->vbs.push_back(vb);
        g->ibs.push_back(ib);
        luzrt_blas* b = nullptr;
        if (luzrt_blas_build(ctx, vb, sizeof(MeshVertex), (uint32_t)m.vertices.size(),
                             (const uint32_t*)ib, (uint32_t)m.indices.size(), &b) != LUZRT_OK) {
            luzgpu_destroy(g);
            return nullptr;
        }
        g->blas.push_back(b);
    }
    uint32_t n = (uint32_t)s->meshNodes.size();
    if (luzrt_tlas_create(ctx, n ? n : 1, &g->tlas) != LUZRT_OK || luzgpu_update(g, s) != LUZRT_OK) {
        luzgpu_destroy(g);
        return nullptr;
    }
    return g;
}

uint32_t luzgpu_instances(luzgpu* g, const luzscene* s, luzrt_instance* out) {
    if (!g || !s) return 0;
    // GPUScene::UpdateResourcesGPU -> CmdBuildTLAS (GPUScene.cpp:354-365).
    for (size_t i = 0; i < s->meshNodes.size() && out; ++i) {
        const MeshNode& n = s->meshNodes[i];
        M34 w = n.GetWorldTransform();
        luzrt_instance& d = out[i];
        d.blas = g->blas[n.mesh];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) d.object_to_world[r][c] = (float)w.m[r][c];
        d.custom_index = (uint32_t)i;
        d.mask = 0xFFu;
    }
    return (uint32_t)s->meshNodes.size();
}

luzrt_status luzgpu_update(luzgpu* g, const luzscene* s) {
    if (!g || !s) return LUZRT_E_INVALID_ARG;
    std::vector<luzrt_instance> inst(s->meshNodes.size());
    luzgpu_instances(g, s, inst.data());
    return luzrt_tlas_build(g->ctx, g->tlas, inst.data(), (uint32_t)inst.size());
}

luzrt_tlas* luzgpu_tlas(luzgpu* g) { return g ? g->tlas : nullptr; }
luzrt_blas* luzgpu_blas(luzgpu* g, uint32_t i) { return g && i < g->blas.size() ? g->blas[i] : nullptr; }
const void* luzgpu_vertices(luzgpu* g, uint32_t i) { return g && i < g->vbs.size() ? g->vbs[i] : nullptr; }

void luzgpu_destroy(luzgpu* g) {
    if (!g) return;
    if (g->tlas) luzrt_tlas_destroy(g->ctx, g->tlas);
    for (auto* b : g->blas) luzrt_blas_destroy(g->ctx, b);
    for (void* p : g->vbs) luzrt_free(g->ctx, p);
    for (void* p : g->ibs) luzrt_free(g->ctx, p);
    delete g;
}

}  // extern "C"
