// scene_model.hpp — Luz's scene model on the host (C++17), shared by the
// generators and GPUScene packer (scene.cpp) and the importers (ingest.cpp).
//
// Model (source/Resources/AssetManager.hpp): MeshAsset/MeshVertex :82-97,
// MaterialAsset :99-112, Node :114-174, MeshNode :176-182, LightNode
// :184-227, CameraNode :229-275, SceneAsset :277-296.  Transforms follow
// Node::ComposeTransform (AssetManager.cpp:38-43: parent * T *
// R(quat(radians(euler))) * S) and glm's column-major conventions; they are
// evaluated in FP64 and rounded once to FP32 where the device reads them.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/luzrt.h"
#include "../../include/luzscene.h"

namespace luz {

constexpr double kPiD = 3.14159265358979323846;

// ------------------------------------------------------------------ RNG
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    uint32_t byte() { return (uint32_t)(next() >> 56); }
};

// Independent sub-stream of a master seed.
enum Stream : uint64_t { kGeometry = 1, kTransforms = 2, kLights = 3, kMaterials = 4, kAnimation = 5,
                         kLayout = 6 };
inline SplitMix64 stream(uint64_t seed, uint64_t id) {
    SplitMix64 m(seed ^ (0xD1B54A32D192ED03ull * id));
    return SplitMix64(m.next());
}

// ------------------------------------------------------------------ math
struct D3 { double x, y, z; };
inline D3 d3(double x, double y, double z) { return D3{x, y, z}; }
inline D3 operator+(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline D3 operator-(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline D3 operator*(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double length(D3 a) { return std::sqrt(dot(a, a)); }
inline D3 normalize(D3 a) { return a * (1.0 / length(a)); }

struct Quat { double w, x, y, z; };

// glm::quat(vec3 eulerAngles) (radians).
inline Quat quat_from_euler(D3 e) {
    double cx = std::cos(e.x * 0.5), cy = std::cos(e.y * 0.5), cz = std::cos(e.z * 0.5);
    double sx = std::sin(e.x * 0.5), sy = std::sin(e.y * 0.5), sz = std::sin(e.z * 0.5);
    return Quat{cx * cy * cz + sx * sy * sz, sx * cy * cz - cx * sy * sz,
                cx * sy * cz + sx * cy * sz, cx * cy * sz - sx * sy * cz};
}

// glm::eulerAngles(quat) (radians): pitch, yaw, roll.
inline D3 euler_from_quat(Quat q) {
    double pitch = std::atan2(2.0 * (q.y * q.z + q.w * q.x), q.w * q.w - q.x * q.x - q.y * q.y + q.z * q.z);
    double s = -2.0 * (q.x * q.z - q.w * q.y);
    s = s < -1.0 ? -1.0 : (s > 1.0 ? 1.0 : s);
    double yaw = std::asin(s);
    double roll = std::atan2(2.0 * (q.x * q.y + q.w * q.z), q.w * q.w + q.x * q.x - q.y * q.y - q.z * q.z);
    return d3(pitch, yaw, roll);
}

// Row-major 3x4 affine transform.
struct M34 {
    double m[3][4];
    static M34 identity() {
        M34 r{};
        for (int i = 0; i < 3; ++i) r.m[i][i] = 1.0;
        return r;
    }
    D3 point(D3 p) const {
        return d3(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
                  m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
                  m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]);
    }
    D3 vec(D3 v) const {
        return d3(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z,
                  m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                  m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
    }
};

inline M34 operator*(const M34& a, const M34& b) {
    M34 r{};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a.m[i][k] * b.m[k][j];
            if (j == 3) s += a.m[i][3];
            r.m[i][j] = s;
        }
    }
    return r;
}

// Rotation matrix of a unit quaternion (glm::mat3_cast, transposed into row-major).
inline M34 quat_matrix(Quat q) {
    double xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    double xz = q.x * q.z, xy = q.x * q.y, yz = q.y * q.z;
    double wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M34 r = M34::identity();
    r.m[0][0] = 1.0 - 2.0 * (yy + zz); r.m[1][0] = 2.0 * (xy + wz);       r.m[2][0] = 2.0 * (xz - wy);
    r.m[0][1] = 2.0 * (xy - wz);       r.m[1][1] = 1.0 - 2.0 * (xx + zz); r.m[2][1] = 2.0 * (yz + wx);
    r.m[0][2] = 2.0 * (xz + wy);       r.m[1][2] = 2.0 * (yz - wx);       r.m[2][2] = 1.0 - 2.0 * (xx + yy);
    return r;
}

// ------------------------------------------------------------------ model
struct MeshVertex {
    float position[3];
    float normal[3];
    float tangent[4];
    float texCoord[2];
};
static_assert(sizeof(MeshVertex) == 48, "MeshVertex is 48 bytes (AssetManager.hpp:83-89)");

struct MeshAsset {
    std::vector<MeshVertex> vertices;
    std::vector<uint32_t> indices;
};

struct MaterialAsset {
    float color[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    float emission[3] = {0.0f, 0.0f, 0.0f};
    float metallic = 0.0f;
    float roughness = 0.5f;
};

struct Node {
    D3 position = d3(0, 0, 0);
    D3 rotation = d3(0, 0, 0);  // Euler angles in degrees
    D3 scale = d3(1, 1, 1);
    M34 parent = M34::identity();   // world transform of the parent chain
    bool has_parent = false;
    // Node::GetWorldTransform = parent * ComposeTransform (AssetManager.cpp:38-59).
    M34 GetWorldTransform() const {
        M34 t = M34::identity();
        t.m[0][3] = position.x; t.m[1][3] = position.y; t.m[2][3] = position.z;
        D3 rad = rotation * (kPiD / 180.0);
        M34 r = rotation_of(rad);
        M34 s = M34::identity();
        s.m[0][0] = scale.x; s.m[1][1] = scale.y; s.m[2][2] = scale.z;
        M34 local = t * r * s;
        return has_parent ? parent * local : local;
    }
    static M34 rotation_of(D3 rad) { return quat_matrix(quat_from_euler(rad)); }
};

struct MeshNode : Node {
    uint32_t mesh = 0;
    uint32_t material = 0;
    // seeded rigid motion (animation)
    D3 base_position = d3(0, 0, 0), base_rotation = d3(0, 0, 0);
    D3 amp = d3(0, 0, 0), spin = d3(0, 0, 0);
    double freq = 0.0, phase = 0.0;
};

struct LightNode : Node {
    enum LightType { Point = 0, Spot = 1, Directional = 2 };
    float color[3] = {1.0f, 1.0f, 1.0f};
    float intensity = 10.0f;
    LightType lightType = Point;
    float radius = 2.0f;
    float innerAngle = 60.f;
    float outerAngle = 50.f;
    float shadowMapRange = 3.0f;    // AssetManager.hpp:208-209
    float shadowMapFar = 2000.0f;
    // Orient the node so that W*(0,-1,0,0) points along d (shortest arc).
    void SetDirection(D3 d) {
        D3 a = d3(0, -1, 0), b = normalize(d);
        double c = dot(a, b);
        Quat q;
        if (c < -1.0 + 1e-12) {
            q = Quat{0.0, 1.0, 0.0, 0.0};  // 180 degrees about x
        } else {
            D3 x = cross(a, b);
            q = Quat{1.0 + c, x.x, x.y, x.z};
            double n = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
            q = Quat{q.w / n, q.x / n, q.y / n, q.z / n};
        }
        rotation = euler_from_quat(q) * (180.0 / kPiD);
    }
};

struct CameraNode {
    D3 eye = d3(0, 0, 0), center = d3(0, 0, 0);
    float horizontalFov = 60.0f;  // passed to perspective() as fovy (AssetManager.cpp:368)
    float nearDistance = 0.01f;
    float farDistance = 1000.0f;
};

}  // namespace luz

struct luzscene {
    std::vector<luz::MeshAsset> meshes;
    std::vector<luz::MaterialAsset> materials;
    std::vector<luz::MeshNode> meshNodes;
    std::vector<luz::LightNode> lights;
    luz::CameraNode camera;
    float ambientLightColor[3] = {1.0f, 1.0f, 1.0f};
    float ambientLight = 0.01f;
    int aoSamples = 4;
    int lightSamples = 2;
    float aoMin = 0.0001f;
    float aoMax = 1.0f;
    int shadowType = LUZRT_SHADOW_RAYTRACED;
    uint32_t width = 256, height = 256;
};

