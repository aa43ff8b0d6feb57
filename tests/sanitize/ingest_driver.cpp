// Host-side driver for the sanitizer test of the scene model and the scene
// importers (built with -fsanitize=address,undefined by
// tests/test_host_sanitizers.py).  Every path on the command line is loaded
// with luzscene_load; a scene that loads is walked through the accessors that
// read its meshes, nodes, lights and camera, saved, and loaded again.  Files
// that fail to load must fail with an error message, never a crash or
// undefined behaviour.  Prints one line per file; exits 0 when every file was
// handled.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/luzrt.h"
#include "../../include/luzscene.h"

namespace {

double walk(const luzscene* s) {
    double acc = 0.0;
    const uint32_t nm = luzscene_mesh_count(s);
    for (uint32_t m = 0; m < nm; ++m) {
        const float* v = nullptr;
        const uint32_t* idx = nullptr;
        uint32_t nv = 0, ni = 0;
        if (luzscene_mesh(s, m, &v, &nv, &idx, &ni) != 0) continue;
        for (uint32_t i = 0; i < nv * 12; ++i) acc += v[i];
        for (uint32_t i = 0; i < ni; ++i) acc += idx[i];
    }
    const uint32_t nn = luzscene_node_count(s);
    std::vector<float> o2w((size_t)nn * 12 + 1);
    std::vector<uint32_t> mesh((size_t)nn + 1);
    std::vector<luzrt_material> mats((size_t)nn + 1);
    if (luzscene_nodes(s, o2w.data(), mesh.data(), mats.data()) == 0)
        for (uint32_t i = 0; i < nn; ++i) acc += o2w[(size_t)i * 12] + mesh[i];
    luzrt_scene sb;
    std::memset(&sb, 0, sizeof sb);
    if (luzscene_pack(s, &sb) == 0) {
        acc += sb.num_lights;
        for (int l = 0; l < sb.num_lights; ++l) {
            luzrt_shadow_map sm;
            if (luzscene_shadow_map(s, (uint32_t)l, 64, &sm) == 0) acc += sm.resolution;
        }
    }
    uint32_t w = 0, h = 0;
    luzrt_camera cam;
    if (luzscene_camera(s, &w, &h, &cam) == 0) acc += w + h;
    float vp[16], ivp[16];
    if (luzscene_view_proj(s, nullptr, vp, ivp) == 0) acc += vp[0];
    acc += (double)luzscene_instanced_triangles(s);
    return acc;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <scratch dir> <scene files...>\n", argv[0]);
        return 2;
    }
    const std::string scratch = argv[1];
    for (int i = 2; i < argc; ++i) {
        char err[256] = {0};
        uint32_t missing = 0;
        luzscene* s = luzscene_load(argv[i], nullptr, &missing, err, sizeof err);
        if (!s) {
            if (err[0] == 0) {
                std::printf("%s: failed without an error message\n", argv[i]);
                return 1;
            }
            std::printf("%s: rejected: %s\n", argv[i], err);
            continue;
        }
        const double a = walk(s);
        const std::string out = scratch + "/resave.luz", bin = scratch + "/resave.luzbin";
        char err2[256] = {0};
        if (luzscene_save(s, out.c_str(), bin.c_str(), err2, sizeof err2) != 0) {
            std::printf("%s: save failed: %s\n", argv[i], err2);
            luzscene_destroy(s);
            return 1;
        }
        luzscene* r = luzscene_load(out.c_str(), bin.c_str(), &missing, err2, sizeof err2);
        if (!r) {
            std::printf("%s: reload failed: %s\n", argv[i], err2);
            luzscene_destroy(s);
            return 1;
        }
        const double b = walk(r);
        std::printf("%s: loaded (%u meshes, %u nodes, %s round trip)\n", argv[i], luzscene_mesh_count(s),
                    luzscene_node_count(s), a == b ? "same" : "different");
        luzscene_destroy(r);
        luzscene_destroy(s);
    }
    return 0;
}
