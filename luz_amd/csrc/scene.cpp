// scene.cpp — the workload generators and the GPUScene packer over the
// luzrt C ABI, on the host scene model of scene_model.hpp (the reference
// rounds every glm step to FP32; this model evaluates transforms in FP64
// and rounds once).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "../../include/luzrt.h"
#include "../../include/luzscene.h"

#include "scene_model.hpp"

namespace luz {
namespace {

// ------------------------------------------------------------------ meshes
void finish_normals(MeshAsset& m) {
    std::vector<D3> acc(m.vertices.size(), d3(0, 0, 0));
    auto P = [&](uint32_t i) { const float* p = m.vertices[i].position; return d3(p[0], p[1], p[2]); };
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
        uint32_t a = m.indices[t], b = m.indices[t + 1], c = m.indices[t + 2];
        D3 n = cross(P(b) - P(a), P(c) - P(a));  // area weighted
        acc[a] = acc[a] + n; acc[b] = acc[b] + n; acc[c] = acc[c] + n;
    }
    for (size_t i = 0; i < m.vertices.size(); ++i) {
        double l = length(acc[i]);
        D3 n = l > 0 ? acc[i] * (1.0 / l) : d3(0, 1, 0);
        m.vertices[i].normal[0] = (float)n.x;
        m.vertices[i].normal[1] = (float)n.y;
        m.vertices[i].normal[2] = (float)n.z;
    }
}

MeshVertex vert(D3 p, D3 n) {
    MeshVertex v{};
    v.position[0] = (float)p.x; v.position[1] = (float)p.y; v.position[2] = (float)p.z;
    v.normal[0] = (float)n.x; v.normal[1] = (float)n.y; v.normal[2] = (float)n.z;
    v.tangent[0] = 1.0f; v.tangent[3] = 1.0f;
    return v;
}

// Axis-aligned box with 4 vertices per face (24 vertices, 12 triangles), like assets/cube.glb.
MeshAsset make_cube(D3 center, double half) {
    MeshAsset m;
    const D3 axes[3] = {d3(1, 0, 0), d3(0, 1, 0), d3(0, 0, 1)};
    for (int a = 0; a < 3; ++a) {
        for (int sgn = -1; sgn <= 1; sgn += 2) {
            D3 n = axes[a] * (double)sgn;
            D3 u = axes[(a + 1) % 3], v = axes[(a + 2) % 3];
            if (sgn < 0) std::swap(u, v);
            uint32_t base = (uint32_t)m.vertices.size();
            D3 c = center + n * half;
            m.vertices.push_back(vert(c - u * half - v * half, n));
            m.vertices.push_back(vert(c + u * half - v * half, n));
            m.vertices.push_back(vert(c + u * half + v * half, n));
            m.vertices.push_back(vert(c - u * half + v * half, n));
            uint32_t q[6] = {0, 1, 2, 0, 2, 3};
            for (uint32_t k : q) m.indices.push_back(base + k);
        }
    }
    return m;
}

// Regular grid on y = height spanning [-extent, extent]^2 with nx*nz quads.
MeshAsset make_grid(uint32_t nx, uint32_t nz, double extent, double height) {
    MeshAsset m;
    for (uint32_t j = 0; j <= nz; ++j)
        for (uint32_t i = 0; i <= nx; ++i) {
            double x = -extent + 2.0 * extent * i / nx, z = -extent + 2.0 * extent * j / nz;
            m.vertices.push_back(vert(d3(x, height, z), d3(0, 1, 0)));
        }
    for (uint32_t j = 0; j < nz; ++j)
        for (uint32_t i = 0; i < nx; ++i) {
            uint32_t a = j * (nx + 1) + i, b = a + 1, c = a + (nx + 1), d = c + 1;
            uint32_t q[6] = {a, c, b, b, c, d};
            for (uint32_t k : q) m.indices.push_back(k);
        }
    return m;
}

// 3D value noise (quintic-faded trilinear interpolation of hashed lattice values).
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
    uint64_t h = (uint64_t)x * 0x8CB92BA72F3D8DD7ull ^ (uint64_t)y * 0xD6E8FEB86659FD93ull ^
                 (uint64_t)z * 0xA0761D6478BD642Full ^ seed * 0xE7037ED1A0B428DBull;
    SplitMix64 r(h);
    return r.uniform();
}

double value_noise(D3 p, uint64_t seed) {
    double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    auto fade = [](double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); };
    double ux = fade(tx), uy = fade(ty), uz = fade(tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    auto lerp = [](double a, double b, double t) { return a + (b - a) * t; };
    double x00 = lerp(c[0][0][0], c[1][0][0], ux), x10 = lerp(c[0][1][0], c[1][1][0], ux);
    double x01 = lerp(c[0][0][1], c[1][0][1], ux), x11 = lerp(c[0][1][1], c[1][1][1], ux);
    return lerp(lerp(x00, x10, uy), lerp(x01, x11, uy), uz);
}

// fBm in [-1, 1]: 4 octaves, lacunarity 2, gain 0.5.
double fbm(D3 p, uint64_t seed, int octaves = 4) {
    double sum = 0.0, amp = 1.0, norm = 0.0, f = 1.0;
    for (int o = 0; o < octaves; ++o) {
        sum += amp * (2.0 * value_noise(p * f, seed + (uint64_t)o * 7919u) - 1.0);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

// Level-L icosphere (20 * 4^L triangles) radially displaced by 1 + amp * fBm(2p).
MeshAsset make_icosphere(int level, double amp, uint64_t seed) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<D3> v = {d3(-1, t, 0), d3(1, t, 0), d3(-1, -t, 0), d3(1, -t, 0),
                         d3(0, -1, t), d3(0, 1, t), d3(0, -1, -t), d3(0, 1, -t),
                         d3(t, 0, -1), d3(t, 0, 1), d3(-t, 0, -1), d3(-t, 0, 1)};
    for (auto& p : v) p = normalize(p);
    std::vector<std::array<uint32_t, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            v.push_back(normalize((v[a] + v[b]) * 0.5));
            uint32_t id = (uint32_t)v.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(f.size() * 4);
        for (auto& tri : f) {
            uint32_t a = midpoint(tri[0], tri[1]), b = midpoint(tri[1], tri[2]), c = midpoint(tri[2], tri[0]);
            nf.push_back({tri[0], a, c});
            nf.push_back({tri[1], b, a});
            nf.push_back({tri[2], c, b});
            nf.push_back({a, b, c});
        }
        f.swap(nf);
    }
    MeshAsset m;
    for (auto& p : v) {
        double r = 1.0 + amp * fbm(p * 2.0, seed);
        m.vertices.push_back(vert(p * r, p));
    }
    for (auto& tri : f)
        for (int k = 0; k < 3; ++k) m.indices.push_back(tri[k]);
    finish_normals(m);
    return m;
}

// Heightfield over [-extent, extent]^2 with nx*nz quads, heights amp * fBm.
MeshAsset make_terrain(uint32_t nx, uint32_t nz, double extent, double base, double amp, uint64_t seed) {
    MeshAsset m = make_grid(nx, nz, extent, base);
    for (auto& vx : m.vertices) {
        D3 p = d3(vx.position[0], 0.0, vx.position[2]) * (4.0 / extent);
        vx.position[1] = (float)(base + amp * fbm(d3(p.x, 0.5, p.z), seed, 6));
    }
    finish_normals(m);
    return m;
}

void random_material(SplitMix64& r, MaterialAsset& mat) {
    for (int k = 0; k < 4; ++k) mat.color[k] = (float)r.byte() / 255.0f;
    mat.metallic = (float)r.uniform();
    mat.roughness = (float)r.uniform();
}

Quat random_rotation(SplitMix64& r) {  // Shoemake 1992
    double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
    double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    return Quat{b * std::cos(2 * kPiD * u3), a * std::sin(2 * kPiD * u2), a * std::cos(2 * kPiD * u2),
                b * std::sin(2 * kPiD * u3)};
}

void set_camera(luzscene& s, D3 eye, D3 center, uint32_t w, uint32_t h) {
    s.camera.eye = eye;
    s.camera.center = center;
    s.width = w;
    s.height = h;
}

// Sphere field: `count` instances of `kinds` displaced icospheres of `level`.
void add_sphere_field(luzscene& s, uint64_t seed, int level, uint32_t kinds, uint32_t count,
                      double box, double smin, double smax) {
    SplitMix64 geo = stream(seed, kGeometry), xf = stream(seed, kTransforms),
               mr = stream(seed, kMaterials), an = stream(seed, kAnimation);
    (void)geo;
    uint32_t first_mesh = (uint32_t)s.meshes.size();
    for (uint32_t k = 0; k < kinds; ++k) s.meshes.push_back(make_icosphere(level, 0.1, seed * 131 + k));
    for (uint32_t i = 0; i < count; ++i) {
        MeshNode n;
        n.mesh = first_mesh + i % kinds;
        n.position = d3(xf.uniform(-box, box), xf.uniform(-box, box), xf.uniform(-box, box));
        n.rotation = euler_from_quat(random_rotation(xf)) * (180.0 / kPiD);
        double sc = xf.uniform(smin, smax);
        n.scale = d3(sc, sc, sc);
        n.base_position = n.position;
        n.base_rotation = n.rotation;
        n.amp = d3(an.uniform(0.5, 3.0), an.uniform(0.5, 3.0), an.uniform(0.5, 3.0));
        n.spin = d3(an.uniform(-2, 2), an.uniform(-2, 2), an.uniform(-2, 2));
        n.freq = an.uniform(0.02, 0.1);
        n.phase = an.uniform(0.0, 2 * kPiD);
        MaterialAsset mat;
        random_material(mr, mat);
        n.material = (uint32_t)s.materials.size();
        s.materials.push_back(mat);
        s.meshNodes.push_back(n);
    }
}

void add_static(luzscene& s, MeshAsset m, SplitMix64& mr) {
    s.meshes.push_back(std::move(m));
    MeshNode n;
    n.mesh = (uint32_t)s.meshes.size() - 1;
    MaterialAsset mat;
    random_material(mr, mat);
    n.material = (uint32_t)s.materials.size();
    s.materials.push_back(mat);
    s.meshNodes.push_back(n);
}

void add_point_lights(luzscene& s, SplitMix64& lr, uint32_t n, double box, float intensity,
                      float radius) {
    for (uint32_t i = 0; i < n; ++i) {
        LightNode L;
        L.lightType = LightNode::Point;
        L.position = d3(lr.uniform(-box, box), lr.uniform(-box, box), lr.uniform(-box, box));
        for (int k = 0; k < 3; ++k) L.color[k] = (float)lr.uniform();
        L.intensity = intensity;
        L.radius = radius;
        s.lights.push_back(L);
    }
}

LightNode directional(D3 dir, float intensity, float radius) {
    LightNode L;
    L.lightType = LightNode::Directional;
    L.SetDirection(dir);
    L.position = normalize(dir) * -1000.0;  // sets tMax of the shadow ray (light.frag:140,143)
    L.intensity = intensity;
    L.radius = radius;
    return L;
}

luzscene* gen_cube(uint64_t seed) {
    auto* s = new luzscene();
    SplitMix64 mr = stream(seed, kMaterials);
    add_static(*s, make_grid(1, 1, 10.0, 0.0), mr);           // ground quad, 2 triangles
    add_static(*s, make_cube(d3(0.0, 0.5, 0.0), 0.5), mr);   // unit cube, 12 triangles
    s->lights.push_back(directional(d3(-0.3, -1.0, -0.2), 3.0f, 0.0f));
    set_camera(*s, d3(3, 2, 3), d3(0, 0, 0), 256, 256);
    s->lightSamples = 1;
    s->aoSamples = 4;
    return s;
}

luzscene* gen_spheres(uint64_t seed) {
    auto* s = new luzscene();
    add_sphere_field(*s, seed, 5, 5, 50, 50.0, 0.5, 3.0);
    SplitMix64 mr = stream(seed ^ 0x5151, kMaterials);
    add_static(*s, make_grid(64, 64, 250.0, -55.0), mr);
    SplitMix64 lr = stream(seed, kLights);
    s->lights.push_back(directional(d3(-0.3, -1.0, -0.2), 3.0f, 0.0f));
    add_point_lights(*s, lr, 8, 50.0, 1000.0f, 0.5f);
    set_camera(*s, d3(0, 50, 110), d3(0, -55, -10), 1920, 1080);
    s->lightSamples = 1;
    s->aoSamples = 16;
    return s;
}

luzscene* gen_terrain(uint64_t seed) {
    auto* s = new luzscene();
    // 2048 x 2048 heightfield decimated 2x along z: 2047 x 1023 quads, ~4.19M triangles.
    SplitMix64 mr = stream(seed ^ 0x7777, kMaterials);
    add_static(*s, make_terrain(2047, 1023, 200.0, -20.0, 12.0, seed + 99), mr);
    add_sphere_field(*s, seed, 5, 5, 40, 60.0, 1.0, 4.0);
    for (auto& n : s->meshNodes)
        if (n.mesh != 0) { n.position.y = std::fabs(n.position.y) * 0.5; n.base_position = n.position; }
    SplitMix64 lr = stream(seed, kLights);
    for (int i = 0; i < 4; ++i) {
        D3 d = d3(lr.uniform(-1, 1), -1.0 - lr.uniform(), lr.uniform(-1, 1));
        s->lights.push_back(directional(d, 1.0f, 0.0f));
    }
    add_point_lights(*s, lr, 32, 80.0, 800.0f, 0.5f);
    for (auto& L : s->lights)
        if (L.lightType == LightNode::Point) L.position.y = std::fabs(L.position.y) * 0.5 + 5.0;
    set_camera(*s, d3(0, 60, 180), d3(0, -10, 0), 3840, 2160);
    s->lightSamples = 1;
    s->aoSamples = 8;
    return s;
}

luzscene* gen_random(uint64_t seed, uint32_t tri_budget, uint32_t w, uint32_t h) {
    SplitMix64 lay = stream(seed, kLayout);
    if (tri_budget == 0) tri_budget = (uint32_t)std::exp(std::log(1e3) + lay.uniform() * (std::log(5e5) - std::log(1e3)));
    if (w == 0 || h == 0) {
        const uint32_t res[4][2] = {{320, 240}, {640, 480}, {1280, 720}, {1920, 1080}};
        uint32_t k = (uint32_t)(lay.uniform() * 4.0);
        if (k > 3) k = 3;
        w = res[k][0];
        h = res[k][1];
    }
    auto* s = new luzscene();
    // The 1080p sphere field's generator (gen_spheres) at a triangle budget:
    // displaced icospheres (up to 5 distinct meshes, one BLAS each) at seeded
    // positions in the same 100^3 box, uniform rotations, scales in [0.5, 3],
    // over the same ground (a 64x64-quad grid, 16x16 below 32k triangles), with
    // the icosphere level and instance count filling the rest of the budget;
    // 1 directional + 8 point lights, 1 shadow ray per light + 16 AO rays.
    SplitMix64 mr = stream(seed ^ 0x3131, kMaterials);
    const uint32_t grid = tri_budget >= 32768 ? 64 : 16;
    const uint32_t ground = 2 * grid * grid;
    uint32_t rest = tri_budget > ground ? tri_budget - ground : 0;
    int level = rest >= 20 * 20480 ? 5 : (rest >= 20 * 5120 ? 4 : (rest >= 20 * 1280 ? 3 : 2));
    uint32_t per = 20u << (2 * level);
    uint32_t count = std::max(1u, rest / per);
    add_sphere_field(*s, seed, level, std::min(count, 5u), count, 50.0, 0.5, 3.0);
    add_static(*s, make_grid(grid, grid, 250.0, -55.0), mr);
    SplitMix64 lr = stream(seed, kLights);
    s->lights.push_back(directional(d3(-0.3, -1.0, -0.2), 3.0f, 0.0f));
    add_point_lights(*s, lr, 8, 50.0, 1000.0f, 0.5f);
    set_camera(*s, d3(0, 50, 110), d3(0, -55, -10), w, h);
    s->lightSamples = 1;
    s->aoSamples = 16;
    return s;
}

inline uint32_t unorm8(double x) {
    x = x < 0 ? 0 : (x > 1 ? 1 : x);
    return (uint32_t)std::lround(x * 255.0);
}

}  // namespace
}  // namespace luz

using namespace luz;

extern "C" {

luzscene* luzscene_generate(int kind, uint64_t seed, uint32_t tri_budget, uint32_t w, uint32_t h) {
    switch (kind) {
        case LUZSCENE_CUBE: return gen_cube(seed);
        case LUZSCENE_SPHERES: return gen_spheres(seed);
        case LUZSCENE_TERRAIN: return gen_terrain(seed);
        case LUZSCENE_RANDOM: return gen_random(seed, tri_budget, w, h);
        default: return nullptr;
    }
}

void luzscene_destroy(luzscene* s) { delete s; }

uint32_t luzscene_mesh_count(const luzscene* s) { return s ? (uint32_t)s->meshes.size() : 0u; }

int luzscene_mesh(const luzscene* s, uint32_t i, const float** v, uint32_t* nv, const uint32_t** idx,
                  uint32_t* ni) {
    if (!s || i >= s->meshes.size()) return -1;
    const MeshAsset& m = s->meshes[i];
    *v = reinterpret_cast<const float*>(m.vertices.data());
    *nv = (uint32_t)m.vertices.size();
    *idx = m.indices.data();
    *ni = (uint32_t)m.indices.size();
    return 0;
}

uint32_t luzscene_node_count(const luzscene* s) { return s ? (uint32_t)s->meshNodes.size() : 0u; }

uint64_t luzscene_instanced_triangles(const luzscene* s) {
    uint64_t t = 0;
    if (!s) return 0;
    for (const auto& n : s->meshNodes) t += s->meshes[n.mesh].indices.size() / 3;
    return t;
}

int luzscene_nodes(const luzscene* s, float* o2w, uint32_t* mesh_ids, luzrt_material* mats) {
    if (!s) return -1;
    for (size_t i = 0; i < s->meshNodes.size(); ++i) {
        const MeshNode& n = s->meshNodes[i];
        M34 w = n.GetWorldTransform();
        if (o2w)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) o2w[i * 12 + r * 4 + c] = (float)w.m[r][c];
        if (mesh_ids) mesh_ids[i] = n.mesh;
        if (mats) {
            const MaterialAsset& m = s->materials[n.material];
            luzrt_material e{};
            e.albedo = unorm8(m.color[0]) | unorm8(m.color[1]) << 8 | unorm8(m.color[2]) << 16 |
                       unorm8(m.color[3]) << 24;
            // opaque.frag:57: (roughness, metallic, occlusion = 1, 1)
            e.material = unorm8(m.roughness) | unorm8(m.metallic) << 8 | 255u << 16 | 255u << 24;
            e.emission = unorm8(m.emission[0]) | unorm8(m.emission[1]) << 8 |
                         unorm8(m.emission[2]) << 16 | 255u << 24;
            mats[i] = e;
        }
    }
    return 0;
}

int luzscene_pack(const luzscene* s, luzrt_scene* out) {
    if (!s || !out) return -1;
    if (s->lights.size() > LUZRT_MAX_LIGHTS) return -2;
    std::memset(out, 0, sizeof(*out));
    // GPUScene::UpdateResources (GPUScene.cpp:239-341)
    for (size_t i = 0; i < s->lights.size(); ++i) {
        const LightNode& L = s->lights[i];
        luzrt_light& b = out->lights[i];
        for (int k = 0; k < 3; ++k) b.color[k] = L.color[k];
        b.intensity = L.intensity;
        D3 wp = L.GetWorldTransform().point(d3(0, 0, 0));   // Node::GetWorldPosition
        b.position[0] = (float)wp.x;
        b.position[1] = (float)wp.y;
        b.position[2] = (float)wp.z;
        b.inner_angle = (float)(L.innerAngle * kPiD / 180.0);
        D3 dir = L.GetWorldTransform().vec(d3(0, -1, 0));
        b.direction[0] = (float)dir.x;
        b.direction[1] = (float)dir.y;
        b.direction[2] = (float)dir.z;
        b.outer_angle = (float)(L.outerAngle * kPiD / 180.0);
        b.type = (int32_t)L.lightType;
        b.num_shadow_samples = s->shadowType == LUZRT_SHADOW_RAYTRACED ? s->lightSamples : 0;
        b.radius = L.radius;
    }
    out->num_lights = (int32_t)s->lights.size();
    for (int k = 0; k < 3; ++k) out->ambient_color[k] = s->ambientLightColor[k];
    out->ambient_intensity = s->ambientLight;
    out->cam_pos[0] = (float)s->camera.eye.x;
    out->cam_pos[1] = (float)s->camera.eye.y;
    out->cam_pos[2] = (float)s->camera.eye.z;
    out->ao_min = s->aoMin;
    out->ao_max = s->aoMax;
    out->ao_num_samples = s->aoSamples > 0 ? s->aoSamples : 0;
    out->frame = 0;
    out->shadow_type = s->shadowType;
    return 0;
}

int luzscene_camera(const luzscene* s, uint32_t* w, uint32_t* h, luzrt_camera* cam) {
    if (!s) return -1;
    if (w) *w = s->width;
    if (h) *h = s->height;
    if (cam) {
        // glm::lookAt(eye, center, +Y) and perspective(radians(fov) as fovy, W/H) with the
        // Vulkan y flip (AssetManager.cpp:334-377).
        D3 f = normalize(s->camera.center - s->camera.eye);
        D3 r = normalize(cross(f, d3(0, 1, 0)));
        D3 u = cross(r, f);
        double th = std::tan(s->camera.horizontalFov * kPiD / 180.0 * 0.5);
        double aspect = (double)s->width / (double)s->height;
        D3 R = r * (th * aspect), U = u * (-th);
        const D3 e = s->camera.eye;
        cam->eye[0] = (float)e.x; cam->eye[1] = (float)e.y; cam->eye[2] = (float)e.z;
        cam->forward[0] = (float)f.x; cam->forward[1] = (float)f.y; cam->forward[2] = (float)f.z;
        cam->right[0] = (float)R.x; cam->right[1] = (float)R.y; cam->right[2] = (float)R.z;
        cam->up[0] = (float)U.x; cam->up[1] = (float)U.y; cam->up[2] = (float)U.z;
        cam->tmin = 0.0f;
        cam->tmax = s->camera.farDistance;
    }
    return 0;
}

}  // extern "C"

namespace luz {
namespace {

// glm::lookAt (right-handed), column-major.
void look_at(D3 eye, D3 center, D3 up, double V[16]) {
    D3 f = normalize(center - eye);
    D3 r = normalize(cross(f, up));
    D3 u = cross(r, f);
    const double v[16] = {r.x, u.x, -f.x, 0, r.y, u.y, -f.y, 0, r.z, u.z, -f.z, 0,
                          -dot(r, eye), -dot(u, eye), dot(f, eye), 1};
    std::memcpy(V, v, sizeof v);
}

void mat4_mul(const double A[16], const double B[16], double C[16]) {
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += A[k * 4 + row] * B[col * 4 + k];
            C[col * 4 + row] = acc;
        }
}

// Gauss-Jordan inverse with partial pivoting; false when singular.
bool mat4_inverse(const double M[16], double out[16]) {
    double a[4][8];
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) {
            a[row][col] = M[col * 4 + row];
            a[row][4 + col] = row == col ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int row = col + 1; row < 4; ++row)
            if (std::fabs(a[row][col]) > std::fabs(a[piv][col])) piv = row;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < 8; ++k) std::swap(a[piv][k], a[col][k]);
        double d = a[col][col];
        for (int k = 0; k < 8; ++k) a[col][k] /= d;
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            double m = a[row][col];
            for (int k = 0; k < 8; ++k) a[row][k] -= m * a[col][k];
        }
    }
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) out[col * 4 + row] = a[row][4 + col];
    return true;
}

// CameraNode::GetProj(zNear, zFar): perspective(radians(horizontalFov) as fovy,
// W/H) in [0, 1] depth with proj[1][1] negated (AssetManager.cpp:363-377).
void camera_proj(const luzscene* s, double n, double fa, double P[16]) {
    const double th = std::tan(s->camera.horizontalFov * kPiD / 180.0 * 0.5);
    const double aspect = (double)s->width / (double)s->height;
    for (int i = 0; i < 16; ++i) P[i] = 0.0;
    P[0] = 1.0 / (aspect * th);
    P[5] = -1.0 / th;
    P[10] = fa / (n - fa);
    P[11] = -1.0;
    P[14] = -(fa * n) / (fa - n);
}

}  // namespace
}  // namespace luz

extern "C" {

int luzscene_shadow_map(const luzscene* s, uint32_t light, uint32_t resolution, luzrt_shadow_map* out) {
    if (!s || !out || light >= s->lights.size() || resolution == 0) return -1;
    const LightNode& L = s->lights[light];
    std::memset(out, 0, sizeof(*out));
    out->light_type = (int32_t)L.lightType;
    out->resolution = resolution;
    M34 W = L.GetWorldTransform();
    D3 pos = W.point(d3(0, 0, 0));
    out->light_pos[0] = (float)pos.x; out->light_pos[1] = (float)pos.y; out->light_pos[2] = (float)pos.z;
    out->z_far = L.shadowMapFar;
    for (int i = 0; i < 16; ++i) out->view_proj[i] = out->inv_view_proj[i] = (i % 5 == 0) ? 1.0f : 0.0f;
    if (L.lightType == LightNode::Point) return 0;   // cube faces use the light position and z_far
    // GPUScene::UpdateResources (GPUScene.cpp:278-311): an orthographic box
    // around the camera frustum cut at farDistance / shadowMapRange, seen
    // from the light's front direction.
    const CameraNode& c = s->camera;
    double V[16], P[16], PV[16], inv[16];
    look_at(c.eye, c.center, d3(0, 1, 0), V);
    camera_proj(s, c.nearDistance, (double)c.farDistance / L.shadowMapRange, P);
    mat4_mul(P, V, PV);
    if (!mat4_inverse(PV, inv)) return -2;
    D3 corners[8];
    D3 center = d3(0, 0, 0);
    int k8 = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double x = 2.0 * i - 1.0, y = 2.0 * j - 1.0, z = 2.0 * k - 1.0, q[4];
                for (int r = 0; r < 4; ++r) q[r] = inv[r] * x + inv[4 + r] * y + inv[8 + r] * z + inv[12 + r];
                corners[k8] = d3(q[0] / q[3], q[1] / q[3], q[2] / q[3]);
                center = center + corners[k8];
                ++k8;
            }
    center = center * (1.0 / 8.0);
    D3 front = W.vec(d3(0, -1, 0));   // Node::GetWorldFront
    double LV[16];
    look_at(center + front, center, d3(0, 1, 0), LV);
    D3 lo = d3(1e300, 1e300, 1e300), hi = d3(-1e300, -1e300, -1e300);
    for (const D3& p : corners) {
        D3 v = d3(LV[0] * p.x + LV[4] * p.y + LV[8] * p.z + LV[12], LV[1] * p.x + LV[5] * p.y + LV[9] * p.z + LV[13],
                  LV[2] * p.x + LV[6] * p.y + LV[10] * p.z + LV[14]);
        lo = d3(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
        hi = d3(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
    }
    const double r = L.shadowMapRange;
    lo.z = lo.z < 0 ? lo.z * r : lo.z / r;
    hi.z = hi.z < 0 ? hi.z / r : hi.z * r;
    // glm::ortho(lo.x, hi.x, lo.y, hi.y, zNear = hi.z, zFar = lo.z), [0, 1] depth
    double O[16] = {0};
    const double zn = hi.z, zf = lo.z;
    O[0] = 2.0 / (hi.x - lo.x);
    O[5] = 2.0 / (hi.y - lo.y);
    O[10] = -1.0 / (zf - zn);
    O[12] = -(hi.x + lo.x) / (hi.x - lo.x);
    O[13] = -(hi.y + lo.y) / (hi.y - lo.y);
    O[14] = -zn / (zf - zn);
    O[15] = 1.0;
    double VP[16], IVP[16];
    mat4_mul(O, LV, VP);
    if (!mat4_inverse(VP, IVP)) return -2;
    for (int i = 0; i < 16; ++i) {
        out->view_proj[i] = (float)VP[i];
        out->inv_view_proj[i] = (float)IVP[i];
    }
    return 0;
}

}  // extern "C"

namespace {

// CameraNode::GetView (glm::lookAt, right-handed) and GetProjJittered =
// translate(jitter) * perspective(radians(horizontalFov) as fovy, W/H, near,
// far) in [0, 1] depth with proj[1][1] negated (AssetManager.cpp:334-377),
// column-major FP64.
void camera_view_proj(const luzscene* s, const float jitter[2], double V[16], double P[16]) {
    const CameraNode& c = s->camera;
    D3 f = normalize(c.center - c.eye);
    D3 r = normalize(cross(f, d3(0, 1, 0)));
    D3 u = cross(r, f);
    const double v[16] = {r.x, u.x, -f.x, 0, r.y, u.y, -f.y, 0, r.z, u.z, -f.z, 0,
                          -dot(r, c.eye), -dot(u, c.eye), dot(f, c.eye), 1};
    for (int i = 0; i < 16; ++i) V[i] = v[i];
    const double th = std::tan(c.horizontalFov * kPiD / 180.0 * 0.5);
    const double aspect = (double)s->width / (double)s->height;
    const double n = c.nearDistance, fa = c.farDistance;
    for (int i = 0; i < 16; ++i) P[i] = 0.0;
    P[0] = 1.0 / (aspect * th);
    P[5] = -1.0 / th;
    P[10] = fa / (n - fa);
    P[11] = -1.0;
    P[14] = -(fa * n) / (fa - n);
    if (jitter) {   // translate(jx, jy, 0) * P adds jitter * row 3 to rows 0 and 1
        for (int col = 0; col < 4; ++col) {
            P[col * 4 + 0] += jitter[0] * P[col * 4 + 3];
            P[col * 4 + 1] += jitter[1] * P[col * 4 + 3];
        }
    }
}

// Inverse of a column-major 4x4 FP64 matrix by Gauss-Jordan with partial
// pivoting, rounded once to FP32; false when singular.
bool invert4(const double M[16], float out[16]) {
    double a[4][8];
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) {
            a[row][col] = M[col * 4 + row];
            a[row][4 + col] = row == col ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int row = col + 1; row < 4; ++row)
            if (std::fabs(a[row][col]) > std::fabs(a[piv][col])) piv = row;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < 8; ++k) std::swap(a[piv][k], a[col][k]);
        double d = a[col][col];
        for (int k = 0; k < 8; ++k) a[col][k] /= d;
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            double m = a[row][col];
            for (int k = 0; k < 8; ++k) a[row][k] -= m * a[col][k];
        }
    }
    for (int row = 0; row < 4; ++row)
        for (int col = 0; col < 4; ++col) out[col * 4 + row] = (float)a[row][4 + col];
    return true;
}

}  // namespace

extern "C" {

int luzscene_view_proj(const luzscene* s, const float jitter[2], float view_proj[16], float inv_view_proj[16]) {
    if (!s || !view_proj) return -1;
    double V[16], P[16];
    camera_view_proj(s, jitter, V, P);
    double M[16];
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc += P[k * 4 + row] * V[col * 4 + k];
            M[col * 4 + row] = acc;
        }
    for (int i = 0; i < 16; ++i) view_proj[i] = (float)M[i];
    if (inv_view_proj && !invert4(M, inv_view_proj)) return -2;
    return 0;
}

int luzscene_camera_matrices(const luzscene* s, const float jitter[2], float view[16], float proj[16],
                             float inv_view[16], float inv_proj[16]) {
    if (!s) return -1;
    double V[16], P[16];
    camera_view_proj(s, jitter, V, P);
    for (int i = 0; i < 16; ++i) {
        if (view) view[i] = (float)V[i];
        if (proj) proj[i] = (float)P[i];
    }
    if (inv_view && !invert4(V, inv_view)) return -2;
    if (inv_proj && !invert4(P, inv_proj)) return -2;
    return 0;
}

void luzscene_jitter(uint32_t index, uint32_t width, uint32_t height, float jitter[2]) {
    // CameraNode::NextJitter (AssetManager.cpp:383-392): Halton(2, 3) of index % 16 + 1,
    // mapped to [-1, 1] and divided by the extent.
    auto halton = [](uint32_t i, uint32_t b) {
        double f = 1.0, r = 0.0;
        while (i > 0) {
            f /= b;
            r += f * (i % b);
            i /= b;
        }
        return r;
    };
    uint32_t k = index % 16 + 1;
    jitter[0] = (float)((2.0 * halton(k, 2) - 1.0) / width);
    jitter[1] = (float)((2.0 * halton(k, 3) - 1.0) / height);
}

int luzscene_set_shadow_type(luzscene* s, int type) {
    if (!s || type < LUZRT_SHADOW_DISABLED || type > LUZRT_SHADOW_MAP) return -1;
    s->shadowType = type;
    return 0;
}

int luzscene_animate(luzscene* s, uint32_t frame) {
    if (!s) return -1;
    for (auto& n : s->meshNodes) {
        if (n.freq == 0.0) continue;
        double a = n.freq * frame + n.phase;
        n.position = n.base_position + d3(n.amp.x * std::sin(a), n.amp.y * std::sin(1.3 * a),
                                          n.amp.z * std::cos(a));
        n.rotation = n.base_rotation + n.spin * (double)frame;
    }
    return 0;
}

}  // extern "C"
