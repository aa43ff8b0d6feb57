// ingest.cpp — scene ingest into Luz's scene model: the project format
// (.luz JSON + .luzbin blob, Serializer.hpp:34-158, AssetManager.cpp:182-257)
// in both directions, and glTF 2.0 (.gltf/.glb) and Wavefront OBJ (+ MTL)
// import (AssetIO.cpp:117-528).  The reference reads JSON with nlohmann::json,
// glTF with tinygltf and OBJ with tinyobjloader; here a small recursive
// descent JSON reader/writer and the two file formats are written directly.
//
// Semantics kept from the reference:
//   .luz   assets[] (type 1 texture, 2 mesh {vertices, indices} as
//          {offset, size} into .luzbin, 3 material), scenes{uuid: scene},
//          initialScene; scene nodes nest through "children" with types
//          5 Node, 6 MeshNode, 7 LightNode, 8 CameraNode; an orbit camera's
//          eye is recomputed from rotation/zoom/center (CameraNode::GetView).
//   glTF   one mesh per primitive with POSITION/NORMAL/TANGENT/TEXCOORD_0 and
//          u8/u16/u32 indices; pbrMetallicRoughness base colour, metallic and
//          roughness factors and the emissive factor; node TRS (rotation to
//          Euler degrees as glm::eulerAngles) or matrix; children.
//   OBJ    per shape, vertices deduplicated on (position, normal, uv) with
//          v flipped (1 - v); MTL Kd -> colour, Ke -> emission, Pm ->
//          metallic, roughness = 1 - mean(Ks) when Ks != 0 else Pr.
// Deliberate differences: a glTF node's mesh brings all its primitives (the
// reference indexes per-primitive arrays with the mesh index); meshes
// without normals get area-weighted vertex normals; tangents are not
// generated (nothing on the lighting path reads them); textures are not
// loaded (the path uses material factors); an OBJ shape whose faces change
// material becomes one mesh per material run.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "scene_model.hpp"

namespace luz {
namespace {

// ------------------------------------------------------------------ JSON
struct JVal {
    enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
    bool b = false;
    double n = 0.0;
    std::string s;   // string value, or the literal text of a number
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;

    const JVal* get(const char* key) const {
        if (t != Obj) return nullptr;
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num(double dflt = 0.0) const { return t == Num ? n : (t == Bool ? (b ? 1.0 : 0.0) : dflt); }
    uint64_t u64() const {
        if (t != Num) return 0;
        errno = 0;
        char* end = nullptr;
        unsigned long long v = std::strtoull(s.c_str(), &end, 10);
        if (errno == 0 && end && *end == '\0') return v;
        return n < 0 ? 0 : (uint64_t)n;
    }
};

struct JParser {
    const char* p;
    const char* end;
    std::string err;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    bool lit(const char* w) {
        size_t n = std::strlen(w);
        if ((size_t)(end - p) < n || std::strncmp(p, w, n) != 0) return fail("bad literal");
        p += n;
        return true;
    }
    static void utf8(std::string& out, uint32_t c) {
        if (c < 0x80) out += (char)c;
        else if (c < 0x800) { out += (char)(0xC0 | (c >> 6)); out += (char)(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) { out += (char)(0xE0 | (c >> 12)); out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F)); }
        else { out += (char)(0xF0 | (c >> 18)); out += (char)(0x80 | ((c >> 12) & 0x3F)); out += (char)(0x80 | ((c >> 6) & 0x3F)); out += (char)(0x80 | (c & 0x3F)); }
    }
    bool hex4(uint32_t& v) {
        if (end - p < 4) return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool str(std::string& out) {
        if (p >= end || *p != '"') return fail("expected string");
        ++p;
        while (p < end && *p != '"') {
            char c = *p++;
            if (c != '\\') { out += c; continue; }
            if (p >= end) return fail("bad escape");
            char e = *p++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t c1;
                    if (!hex4(c1)) return false;
                    if (c1 >= 0xD800 && c1 < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t c2;
                        if (!hex4(c2)) return false;
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    utf8(out, c1);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool value(JVal& v, int depth) {
        if (depth > 200) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        char c = *p;
        if (c == '{') {
            v.t = JVal::Obj;
            ++p;
            ws();
            if (p < end && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return fail("expected ':'");
                ++p;
                v.o.emplace_back(std::move(k), JVal());
                if (!value(v.o.back().second, depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.t = JVal::Arr;
            ++p;
            ws();
            if (p < end && *p == ']') { ++p; return true; }
            for (;;) {
                v.a.emplace_back();
                if (!value(v.a.back(), depth + 1)) return false;
                ws();
                if (p < end && *p == ',') { ++p; continue; }
                if (p < end && *p == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.t = JVal::Str; return str(v.s); }
        if (c == 't') { v.t = JVal::Bool; v.b = true; return lit("true"); }
        if (c == 'f') { v.t = JVal::Bool; v.b = false; return lit("false"); }
        if (c == 'n') { v.t = JVal::Null; return lit("null"); }
        const char* s0 = p;
        if (*p == '-') ++p;
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) ++p;
        if (p == s0) return fail("unexpected character");
        v.t = JVal::Num;
        v.s.assign(s0, p);
        char* e = nullptr;
        v.n = std::strtod(v.s.c_str(), &e);
        if (!e || *e != '\0') return fail("bad number");
        return true;
    }
};

bool parse_json(const std::string& text, JVal& out, std::string& err) {
    JParser P{text.data(), text.data() + text.size(), {}};
    if (!P.value(out, 0)) { err = "JSON: " + P.err; return false; }
    P.ws();
    if (P.p != P.end) { err = "JSON: trailing characters"; return false; }
    return true;
}

// Minimal JSON writer (numbers: %.17g for doubles, %.9g for floats).
struct JOut {
    std::string s;
    void raw(const char* t) { s += t; }
    void key(const char* k) { s += '"'; s += k; s += "\":"; }
    void str(const std::string& v) {
        s += '"';
        for (char c : v) {
            if (c == '"' || c == '\\') { s += '\\'; s += c; }
            else if ((unsigned char)c < 0x20) { char b[8]; std::snprintf(b, sizeof b, "\\u%04x", c); s += b; }
            else s += c;
        }
        s += '"';
    }
    void d(double v) { char b[40]; std::snprintf(b, sizeof b, "%.17g", v); s += b; }
    void f(float v) { char b[40]; std::snprintf(b, sizeof b, "%.9g", (double)v); s += b; }
    void u(uint64_t v) { char b[32]; std::snprintf(b, sizeof b, "%llu", (unsigned long long)v); s += b; }
    void i(long long v) { char b[32]; std::snprintf(b, sizeof b, "%lld", v); s += b; }
};

bool read_file(const std::string& path, std::string& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

std::string dir_of(const std::string& path) {
    size_t k = path.find_last_of('/');
    return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

std::string ext_of(const std::string& path) {
    size_t k = path.find_last_of('.');
    std::string e = k == std::string::npos ? std::string() : path.substr(k);
    for (char& c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

D3 vec3_of(const JVal* v, D3 dflt) {
    if (!v || v->t != JVal::Arr || v->a.size() != 3) return dflt;
    return d3(v->a[0].num(), v->a[1].num(), v->a[2].num());
}

void area_normals(MeshAsset& m) {
    std::vector<D3> acc(m.vertices.size(), d3(0, 0, 0));
    auto P = [&](uint32_t i) { const float* q = m.vertices[i].position; return d3(q[0], q[1], q[2]); };
    for (size_t t = 0; t + 2 < m.indices.size(); t += 3) {
        uint32_t a = m.indices[t], b = m.indices[t + 1], c = m.indices[t + 2];
        D3 n = cross(P(b) - P(a), P(c) - P(a));
        acc[a] = acc[a] + n; acc[b] = acc[b] + n; acc[c] = acc[c] + n;
    }
    for (size_t i = 0; i < acc.size(); ++i) {
        double l = length(acc[i]);
        D3 n = l > 0 ? acc[i] * (1.0 / l) : d3(0, 1, 0);
        m.vertices[i].normal[0] = (float)n.x; m.vertices[i].normal[1] = (float)n.y; m.vertices[i].normal[2] = (float)n.z;
    }
}

bool valid_mesh(const MeshAsset& m) {
    if (m.indices.size() % 3) return false;
    for (uint32_t i : m.indices)
        if (i >= m.vertices.size()) return false;
    return true;
}

// A camera that frames the scene's instanced bounds from the +x+y+z diagonal.
void frame_camera(luzscene& s) {
    D3 lo = d3(1e300, 1e300, 1e300), hi = d3(-1e300, -1e300, -1e300);
    for (const MeshNode& n : s.meshNodes) {
        M34 w = n.GetWorldTransform();
        for (const MeshVertex& v : s.meshes[n.mesh].vertices) {
            D3 p = w.point(d3(v.position[0], v.position[1], v.position[2]));
            lo = d3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
            hi = d3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
        }
    }
    if (lo.x > hi.x) { lo = d3(-1, -1, -1); hi = d3(1, 1, 1); }
    D3 c = (lo + hi) * 0.5;
    double r = std::max(length(hi - lo) * 0.5, 1e-3);
    s.camera.center = c;
    s.camera.eye = c + normalize(d3(1, 0.8, 1.2)) * (r * 2.5);
    s.camera.farDistance = (float)std::max(1000.0, r * 10.0);
}

// ------------------------------------------------------------------ glTF
struct Gltf {
    JVal j;
    std::vector<std::string> buffers;
};

bool b64_decode(const std::string& in, std::string& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (char c : in) {
        if (c == '=') break;
        int v = val(c);
        if (v < 0) continue;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out += (char)((acc >> bits) & 0xFF);
        }
    }
    return true;
}

struct Accessor {
    const uint8_t* data = nullptr;
    size_t count = 0, stride = 0, comp_size = 0, ncomp = 0;
    int comp_type = 0;
};

bool accessor(const Gltf& g, int index, Accessor& A, std::string& err) {
    const JVal* accs = g.j.get("accessors");
    const JVal* views = g.j.get("bufferViews");
    if (!accs || index < 0 || (size_t)index >= accs->a.size()) { err = "glTF: bad accessor"; return false; }
    const JVal& a = accs->a[index];
    const JVal* bv = a.get("bufferView");
    if (!bv || !views || (size_t)bv->num() >= views->a.size()) { err = "glTF: accessor without buffer view"; return false; }
    const JVal& v = views->a[(size_t)bv->num()];
    size_t buf = (size_t)(v.get("buffer") ? v.get("buffer")->num() : 0);
    if (buf >= g.buffers.size()) { err = "glTF: bad buffer index"; return false; }
    A.comp_type = (int)a.get("componentType")->num();
    A.count = (size_t)a.get("count")->num();
    std::string type = a.get("type") ? a.get("type")->s : "SCALAR";
    A.ncomp = type == "SCALAR" ? 1 : type == "VEC2" ? 2 : type == "VEC3" ? 3 : type == "VEC4" ? 4 : type == "MAT4" ? 16 : 0;
    A.comp_size = A.comp_type == 5126 || A.comp_type == 5125 ? 4 : (A.comp_type == 5123 || A.comp_type == 5122 ? 2 : 1);
    if (A.ncomp == 0) { err = "glTF: unsupported accessor type " + type; return false; }
    size_t off = (size_t)(v.get("byteOffset") ? v.get("byteOffset")->num() : 0) +
                 (size_t)(a.get("byteOffset") ? a.get("byteOffset")->num() : 0);
    A.stride = v.get("byteStride") ? (size_t)v.get("byteStride")->num() : A.comp_size * A.ncomp;
    const std::string& B = g.buffers[buf];
    size_t need = A.count ? off + A.stride * (A.count - 1) + A.comp_size * A.ncomp : off;
    if (need > B.size()) { err = "glTF: accessor outside its buffer"; return false; }
    A.data = reinterpret_cast<const uint8_t*>(B.data()) + off;
    return true;
}

bool read_floats(const Accessor& A, size_t i, float* out, size_t n) {
    if (A.comp_type != 5126) return false;
    std::memcpy(out, A.data + i * A.stride, sizeof(float) * std::min(n, A.ncomp));
    return true;
}

luzscene* load_gltf(const std::string& path, std::string& err) {
    std::string file;
    if (!read_file(path, file)) { err = "cannot read " + path; return nullptr; }
    Gltf g;
    std::string json;
    std::string bin;
    bool glb = file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0;
    if (glb) {
        uint32_t version, length;
        std::memcpy(&version, file.data() + 4, 4);
        std::memcpy(&length, file.data() + 8, 4);
        if (version != 2 || length > file.size()) { err = "GLB: bad header"; return nullptr; }
        size_t off = 12;
        while (off + 8 <= length) {
            uint32_t clen, ctype;
            std::memcpy(&clen, file.data() + off, 4);
            std::memcpy(&ctype, file.data() + off + 4, 4);
            if (off + 8 + (size_t)clen > length) { err = "GLB: chunk past the end"; return nullptr; }
            if (ctype == 0x4E4F534Au) json.assign(file.data() + off + 8, clen);
            else if (ctype == 0x004E4942u) bin.assign(file.data() + off + 8, clen);
            off += 8 + ((clen + 3) & ~3u);
        }
    } else {
        json = file;
    }
    if (!parse_json(json, g.j, err)) return nullptr;
    if (const JVal* bufs = g.j.get("buffers")) {
        for (size_t i = 0; i < bufs->a.size(); ++i) {
            const JVal* uri = bufs->a[i].get("uri");
            std::string data;
            if (!uri) {
                data = bin;   // GLB-stored buffer
            } else if (uri->s.rfind("data:", 0) == 0) {
                size_t comma = uri->s.find(',');
                if (comma == std::string::npos) { err = "glTF: bad data URI"; return nullptr; }
                b64_decode(uri->s.substr(comma + 1), data);
            } else if (!read_file(dir_of(path) + uri->s, data)) {
                err = "glTF: cannot read buffer " + uri->s;
                return nullptr;
            }
            g.buffers.push_back(std::move(data));
        }
    }
    auto* s = new luzscene();
    // materials
    std::vector<int> mat_index;
    if (const JVal* mats = g.j.get("materials")) {
        for (const JVal& m : mats->a) {
            MaterialAsset a;
            if (const JVal* pbr = m.get("pbrMetallicRoughness")) {
                if (const JVal* c = pbr->get("baseColorFactor"))
                    for (size_t k = 0; k < 4 && k < c->a.size(); ++k) a.color[k] = (float)c->a[k].num();
                if (const JVal* v = pbr->get("metallicFactor")) a.metallic = (float)v->num();
                if (const JVal* v = pbr->get("roughnessFactor")) a.roughness = (float)v->num();
            }
            if (const JVal* e = m.get("emissiveFactor"))
                for (size_t k = 0; k < 3 && k < e->a.size(); ++k) a.emission[k] = (float)e->a[k].num();
            s->materials.push_back(a);
        }
    }
    const uint32_t default_mat = (uint32_t)s->materials.size();
    s->materials.push_back(MaterialAsset());
    // meshes: one MeshAsset per primitive
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> mesh_prims;   // (mesh asset, material)
    if (const JVal* meshes = g.j.get("meshes")) {
        for (const JVal& mesh : meshes->a) {
            mesh_prims.emplace_back();
            const JVal* prims = mesh.get("primitives");
            if (!prims) continue;
            for (const JVal& pr : prims->a) {
                if (pr.get("mode") && pr.get("mode")->num() != 4) { err = "glTF: only triangle primitives"; delete s; return nullptr; }
                const JVal* attrs = pr.get("attributes");
                const JVal* pos = attrs ? attrs->get("POSITION") : nullptr;
                if (!pos) { err = "glTF: primitive without POSITION"; delete s; return nullptr; }
                if (!pr.get("indices")) { err = "glTF: non-indexed primitive"; delete s; return nullptr; }
                Accessor P, N, T, UV, I;
                bool hasN = false, hasT = false, hasUV = false;
                if (!accessor(g, (int)pos->num(), P, err)) { delete s; return nullptr; }
                if (const JVal* a = attrs->get("NORMAL")) { if (!accessor(g, (int)a->num(), N, err)) { delete s; return nullptr; } hasN = true; }
                if (const JVal* a = attrs->get("TANGENT")) { if (!accessor(g, (int)a->num(), T, err)) { delete s; return nullptr; } hasT = true; }
                if (const JVal* a = attrs->get("TEXCOORD_0")) { if (!accessor(g, (int)a->num(), UV, err)) { delete s; return nullptr; } hasUV = true; }
                if (!accessor(g, (int)pr.get("indices")->num(), I, err)) { delete s; return nullptr; }
                MeshAsset m;
                m.vertices.resize(P.count);
                for (size_t v = 0; v < P.count; ++v) {
                    MeshVertex& mv = m.vertices[v];
                    std::memset(&mv, 0, sizeof(mv));
                    if (!read_floats(P, v, mv.position, 3)) { err = "glTF: non-float positions"; delete s; return nullptr; }
                    if (hasN && v < N.count) read_floats(N, v, mv.normal, 3);
                    if (hasT && v < T.count) read_floats(T, v, mv.tangent, 4);
                    if (hasUV && v < UV.count) read_floats(UV, v, mv.texCoord, 2);
                }
                m.indices.resize(I.count);
                for (size_t k = 0; k < I.count; ++k) {
                    const uint8_t* q = I.data + k * I.stride;
                    if (I.comp_type == 5125) { uint32_t x; std::memcpy(&x, q, 4); m.indices[k] = x; }
                    else if (I.comp_type == 5123) { uint16_t x; std::memcpy(&x, q, 2); m.indices[k] = x; }
                    else if (I.comp_type == 5121) m.indices[k] = *q;
                    else { err = "glTF: unsupported index type"; delete s; return nullptr; }
                }
                if (!valid_mesh(m)) { err = "glTF: index out of range"; delete s; return nullptr; }
                if (!hasN) area_normals(m);
                int mat = pr.get("material") ? (int)pr.get("material")->num() : -1;
                mesh_prims.back().emplace_back((uint32_t)s->meshes.size(),
                                               mat >= 0 && (size_t)mat < default_mat ? (uint32_t)mat : default_mat);
                s->meshes.push_back(std::move(m));
            }
        }
    }
    // nodes
    const JVal* nodes = g.j.get("nodes");
    size_t nn = nodes ? nodes->a.size() : 0;
    std::vector<Node> local(nn);
    for (size_t i = 0; i < nn; ++i) {
        const JVal& n = nodes->a[i];
        Node& L = local[i];
        if (const JVal* m = n.get("matrix")) {
            if (m->a.size() == 16) {
                // column-major 4x4 -> TRS (no shear or projection, as glm::decompose)
                double c[16];
                for (int k = 0; k < 16; ++k) c[k] = m->a[k].num();
                D3 X = d3(c[0], c[1], c[2]), Y = d3(c[4], c[5], c[6]), Z = d3(c[8], c[9], c[10]);
                double sx = length(X), sy = length(Y), sz = length(Z);
                if (dot(cross(X, Y), Z) < 0) sx = -sx;
                L.position = d3(c[12], c[13], c[14]);
                L.scale = d3(sx, sy, sz);
                double r[3][3] = {{c[0] / sx, c[4] / sy, c[8] / sz}, {c[1] / sx, c[5] / sy, c[9] / sz},
                                  {c[2] / sx, c[6] / sy, c[10] / sz}};
                Quat q;
                double tr = r[0][0] + r[1][1] + r[2][2];
                if (tr > 0) {
                    double k = std::sqrt(tr + 1.0) * 2;
                    q = Quat{0.25 * k, (r[2][1] - r[1][2]) / k, (r[0][2] - r[2][0]) / k, (r[1][0] - r[0][1]) / k};
                } else if (r[0][0] > r[1][1] && r[0][0] > r[2][2]) {
                    double k = std::sqrt(1.0 + r[0][0] - r[1][1] - r[2][2]) * 2;
                    q = Quat{(r[2][1] - r[1][2]) / k, 0.25 * k, (r[0][1] + r[1][0]) / k, (r[0][2] + r[2][0]) / k};
                } else if (r[1][1] > r[2][2]) {
                    double k = std::sqrt(1.0 + r[1][1] - r[0][0] - r[2][2]) * 2;
                    q = Quat{(r[0][2] - r[2][0]) / k, (r[0][1] + r[1][0]) / k, 0.25 * k, (r[1][2] + r[2][1]) / k};
                } else {
                    double k = std::sqrt(1.0 + r[2][2] - r[0][0] - r[1][1]) * 2;
                    q = Quat{(r[1][0] - r[0][1]) / k, (r[0][2] + r[2][0]) / k, (r[1][2] + r[2][1]) / k, 0.25 * k};
                }
                L.rotation = euler_from_quat(q) * (180.0 / kPiD);
            }
        }
        if (const JVal* t = n.get("translation")) L.position = vec3_of(t, L.position);
        if (const JVal* r = n.get("rotation")) {
            if (r->a.size() == 4) {
                Quat q{r->a[3].num(), r->a[0].num(), r->a[1].num(), r->a[2].num()};
                L.rotation = euler_from_quat(q) * (180.0 / kPiD);
            }
        }
        if (const JVal* sc = n.get("scale")) L.scale = vec3_of(sc, L.scale);
    }
    // world transforms from the scene roots down
    std::vector<int> parent(nn, -1);
    for (size_t i = 0; i < nn; ++i)
        if (const JVal* ch = nodes->a[i].get("children"))
            for (const JVal& c : ch->a)
                if ((size_t)c.num() < nn) parent[(size_t)c.num()] = (int)i;
    std::vector<M34> world(nn);
    std::vector<int> state(nn, 0);
    std::function<bool(size_t)> resolve = [&](size_t i) -> bool {
        if (state[i] == 2) return true;
        if (state[i] == 1) return false;   // cycle
        state[i] = 1;
        M34 w = local[i].GetWorldTransform();
        if (parent[i] >= 0) {
            if (!resolve((size_t)parent[i])) return false;
            w = world[(size_t)parent[i]] * w;
        }
        world[i] = w;
        state[i] = 2;
        return true;
    };
    std::vector<size_t> visit;
    int scene_index = g.j.get("scene") ? (int)g.j.get("scene")->num() : 0;
    const JVal* scenes = g.j.get("scenes");
    std::vector<bool> in_scene(nn, scenes == nullptr);
    if (scenes && scene_index >= 0 && (size_t)scene_index < scenes->a.size()) {
        std::vector<size_t> stack;
        if (const JVal* roots = scenes->a[(size_t)scene_index].get("nodes"))
            for (const JVal& r : roots->a) if ((size_t)r.num() < nn) stack.push_back((size_t)r.num());
        while (!stack.empty()) {
            size_t i = stack.back();
            stack.pop_back();
            if (in_scene[i]) continue;
            in_scene[i] = true;
            if (const JVal* ch = nodes->a[i].get("children"))
                for (const JVal& c : ch->a) if ((size_t)c.num() < nn) stack.push_back((size_t)c.num());
        }
    }
    for (size_t i = 0; i < nn; ++i) {
        if (!in_scene[i]) continue;
        if (!resolve(i)) { err = "glTF: node hierarchy has a cycle"; delete s; return nullptr; }
        const JVal& n = nodes->a[i];
        if (const JVal* me = n.get("mesh")) {
            size_t mi = (size_t)me->num();
            if (mi < mesh_prims.size())
                for (auto& pm : mesh_prims[mi]) {
                    MeshNode node;
                    node.mesh = pm.first;
                    node.material = pm.second;
                    node.parent = world[i];
                    node.has_parent = true;
                    s->meshNodes.push_back(node);
                }
        }
        const JVal* ext = n.get("extensions");
        if (ext && ext->get("KHR_lights_punctual")) {
            LightNode L;   // default light parameters, as the reference
            L.parent = world[i];
            L.has_parent = true;
            s->lights.push_back(L);
        }
    }
    frame_camera(*s);
    return s;
}

// ------------------------------------------------------------------ OBJ
struct ObjMtl {
    float Kd[3] = {0, 0, 0}, Ke[3] = {0, 0, 0}, Ks[3] = {0, 0, 0};
    float Pm = 0, Pr = 0;
};

void load_mtl(const std::string& path, std::map<std::string, ObjMtl>& out) {
    std::ifstream f(path);
    std::string line, cur;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "newmtl") { ss >> cur; out[cur] = ObjMtl(); }
        else if (cur.empty()) continue;
        else if (tag == "Kd") ss >> out[cur].Kd[0] >> out[cur].Kd[1] >> out[cur].Kd[2];
        else if (tag == "Ke") ss >> out[cur].Ke[0] >> out[cur].Ke[1] >> out[cur].Ke[2];
        else if (tag == "Ks") ss >> out[cur].Ks[0] >> out[cur].Ks[1] >> out[cur].Ks[2];
        else if (tag == "Pm") ss >> out[cur].Pm;
        else if (tag == "Pr") ss >> out[cur].Pr;
    }
}

struct VertKey {
    float v[8];
    bool operator==(const VertKey& o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
};
struct VertHash {
    size_t operator()(const VertKey& k) const {
        uint64_t h = 1469598103934665603ull;
        const unsigned char* b = reinterpret_cast<const unsigned char*>(k.v);
        for (size_t i = 0; i < sizeof k.v; ++i) h = (h ^ b[i]) * 1099511628211ull;
        return (size_t)h;
    }
};

luzscene* load_obj(const std::string& path, std::string& err) {
    std::ifstream f(path);
    if (!f) { err = "cannot read " + path; return nullptr; }
    std::vector<float> P, N, T;
    std::map<std::string, ObjMtl> mtls;
    auto* s = new luzscene();
    std::map<std::string, uint32_t> mat_id;
    auto material_of = [&](const std::string& name) -> uint32_t {
        auto it = mat_id.find(name);
        if (it != mat_id.end()) return it->second;
        MaterialAsset a;
        auto m = mtls.find(name);
        if (m != mtls.end()) {
            const ObjMtl& o = m->second;
            for (int k = 0; k < 3; ++k) { a.color[k] = o.Kd[k]; a.emission[k] = o.Ke[k]; }
            a.color[3] = 1.0f;
            a.metallic = o.Pm;
            a.roughness = (o.Ks[0] != 0 || o.Ks[1] != 0 || o.Ks[2] != 0) ? 1.0f - (o.Ks[0] + o.Ks[1] + o.Ks[2]) / 3.0f : o.Pr;
        }
        uint32_t id = (uint32_t)s->materials.size();
        s->materials.push_back(a);
        mat_id[name] = id;
        return id;
    };
    const uint32_t no_mat = material_of("");
    MeshAsset cur;
    std::unordered_map<VertKey, uint32_t, VertHash> uniq;
    uint32_t cur_mat = no_mat;
    auto flush = [&]() {
        if (!cur.indices.empty()) {
            bool need_normals = false;
            for (const MeshVertex& v : cur.vertices)
                if (v.normal[0] == 0 && v.normal[1] == 0 && v.normal[2] == 0) need_normals = true;
            if (need_normals) area_normals(cur);
            MeshNode n;
            n.mesh = (uint32_t)s->meshes.size();
            n.material = cur_mat;
            s->meshes.push_back(std::move(cur));
            s->meshNodes.push_back(n);
        }
        cur = MeshAsset();
        uniq.clear();
    };
    std::string line;
    size_t lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") { float x = 0, y = 0, z = 0; ss >> x >> y >> z; P.insert(P.end(), {x, y, z}); }
        else if (tag == "vn") { float x = 0, y = 0, z = 0; ss >> x >> y >> z; N.insert(N.end(), {x, y, z}); }
        else if (tag == "vt") { float u = 0, v = 0; ss >> u >> v; T.insert(T.end(), {u, v}); }
        else if (tag == "o" || tag == "g") flush();
        else if (tag == "mtllib") { std::string m; ss >> m; load_mtl(dir_of(path) + m, mtls); }
        else if (tag == "usemtl") {
            std::string m;
            ss >> m;
            uint32_t id = material_of(m);
            if (id != cur_mat) { flush(); cur_mat = id; }
        } else if (tag == "f") {
            std::vector<uint32_t> poly;
            std::string tok;
            while (ss >> tok) {
                long vi = 0, ti = 0, ni = 0;
                int field = 0;
                std::string part;
                std::istringstream ts(tok);
                while (std::getline(ts, part, '/')) {
                    long v = part.empty() ? 0 : std::strtol(part.c_str(), nullptr, 10);
                    if (field == 0) vi = v; else if (field == 1) ti = v; else ni = v;
                    ++field;
                }
                auto fix = [](long i, size_t n) -> long { return i < 0 ? (long)n + i : i - 1; };
                long pv = fix(vi, P.size() / 3), pt = ti ? fix(ti, T.size() / 2) : -1, pn = ni ? fix(ni, N.size() / 3) : -1;
                if (pv < 0 || (size_t)pv >= P.size() / 3 || (ti && (pt < 0 || (size_t)pt >= T.size() / 2)) ||
                    (ni && (pn < 0 || (size_t)pn >= N.size() / 3))) {
                    err = "OBJ: face index out of range at line " + std::to_string(lineno);
                    delete s;
                    return nullptr;
                }
                VertKey k;
                std::memset(k.v, 0, sizeof k.v);
                for (int c = 0; c < 3; ++c) k.v[c] = P[(size_t)pv * 3 + c];
                if (pn >= 0) for (int c = 0; c < 3; ++c) k.v[3 + c] = N[(size_t)pn * 3 + c];
                if (pt >= 0) { k.v[6] = T[(size_t)pt * 2]; k.v[7] = 1.0f - T[(size_t)pt * 2 + 1]; }
                auto it = uniq.find(k);
                uint32_t idx;
                if (it == uniq.end()) {
                    MeshVertex mv;
                    std::memset(&mv, 0, sizeof mv);
                    std::memcpy(mv.position, k.v, 12);
                    std::memcpy(mv.normal, k.v + 3, 12);
                    mv.texCoord[0] = k.v[6];
                    mv.texCoord[1] = k.v[7];
                    idx = (uint32_t)cur.vertices.size();
                    cur.vertices.push_back(mv);
                    uniq.emplace(k, idx);
                } else {
                    idx = it->second;
                }
                poly.push_back(idx);
            }
            for (size_t k = 2; k < poly.size(); ++k) cur.indices.insert(cur.indices.end(), {poly[0], poly[k - 1], poly[k]});
        }
    }
    flush();
    frame_camera(*s);
    return s;
}

// ------------------------------------------------------------------ .luz / .luzbin
enum ObjType { kTexture = 1, kMesh = 2, kMaterial = 3, kScene = 4, kNode = 5, kMeshNode = 6, kLightNode = 7,
               kCameraNode = 8 };

bool blob_range(const JVal* r, const std::string& bin, size_t& off, size_t& size) {
    if (!r) return false;
    off = (size_t)(r->get("offset") ? r->get("offset")->u64() : 0);
    size = (size_t)(r->get("size") ? r->get("size")->u64() : 0);
    return off <= bin.size() && size <= bin.size() - off;
}

struct LuzLoad {
    luzscene* s;
    const std::string* bin;
    bool have_bin;
    std::map<uint64_t, uint32_t> mesh_of, mat_of;
    uint64_t main_camera = 0;
    uint32_t missing_meshes = 0;
    std::string err;

    void read_node_trs(const JVal& j, Node& n) {
        n.position = vec3_of(j.get("position"), n.position);
        n.rotation = vec3_of(j.get("rotation"), n.rotation);
        n.scale = vec3_of(j.get("scale"), n.scale);
    }

    void node(const JVal& j, const M34& parent, bool has_parent) {
        int type = (int)(j.get("type") ? j.get("type")->num() : 0);
        Node base;
        read_node_trs(j, base);
        base.parent = parent;
        base.has_parent = has_parent;
        M34 world = base.GetWorldTransform();
        if (type == kMeshNode) {
            uint64_t mesh = j.get("mesh") ? j.get("mesh")->u64() : 0;
            uint64_t mat = j.get("material") ? j.get("material")->u64() : 0;
            auto mi = mesh_of.find(mesh);
            if (mi == mesh_of.end()) {
                if (mesh) missing_meshes++;
            } else {
                MeshNode m;
                static_cast<Node&>(m) = base;
                m.mesh = mi->second;
                auto ma = mat_of.find(mat);
                m.material = ma == mat_of.end() ? (uint32_t)(s->materials.size() - 1) : ma->second;
                s->meshNodes.push_back(m);
            }
        } else if (type == kLightNode) {
            LightNode L;
            static_cast<Node&>(L) = base;
            if (const JVal* c = j.get("color")) for (size_t k = 0; k < 3 && k < c->a.size(); ++k) L.color[k] = (float)c->a[k].num();
            if (const JVal* v = j.get("intensity")) L.intensity = (float)v->num();
            if (const JVal* v = j.get("lightType")) L.lightType = (LightNode::LightType)(int)v->num();
            if (const JVal* v = j.get("innerAngle")) L.innerAngle = (float)v->num();
            if (const JVal* v = j.get("outerAngle")) L.outerAngle = (float)v->num();
            if (const JVal* v = j.get("radius")) L.radius = (float)v->num();
            if (const JVal* v = j.get("shadowMapRange")) L.shadowMapRange = (float)v->num();
            if (const JVal* v = j.get("shadowMapFar")) L.shadowMapFar = (float)v->num();
            s->lights.push_back(L);
        } else if (type == kCameraNode) {
            uint64_t uuid = j.get("uuid") ? j.get("uuid")->u64() : 0;
            if (main_camera == 0 || uuid == main_camera) {
                CameraNode& c = s->camera;
                c.center = vec3_of(j.get("center"), c.center);
                c.eye = vec3_of(j.get("eye"), c.eye);
                if (const JVal* v = j.get("horizontalFov")) c.horizontalFov = (float)v->num();
                if (const JVal* v = j.get("nearDistance")) c.nearDistance = (float)v->num();
                if (const JVal* v = j.get("farDistance")) c.farDistance = (float)v->num();
                int mode = (int)(j.get("mode") ? j.get("mode")->num() : 0);
                if (mode == 0) {
                    // orbit camera: CameraNode::GetView recomputes the eye (AssetManager.cpp:336-345)
                    D3 rot = vec3_of(j.get("rotation"), d3(0, 0, 0));
                    double rx = (rot.x + 90.0) * kPiD / 180.0, ry = (rot.y + 90.0) * kPiD / 180.0;
                    D3 dir = d3(std::cos(-ry) * std::sin(rx), std::cos(rx), std::sin(-ry) * std::sin(rx));
                    int ctype = (int)(j.get("cameraType") ? j.get("cameraType")->num() : 0);
                    double zoom = j.get("zoom") ? j.get("zoom")->num() : 1.0;
                    if (ctype == 0) dir = dir * zoom;
                    c.eye = c.center - dir;
                }
            }
        }
        if (const JVal* ch = j.get("children"))
            for (const JVal& c : ch->a) node(c, world, true);
    }
};

luzscene* load_luz(const std::string& path, const std::string& bin_path, std::string& err, uint32_t* missing) {
    std::string text, bin;
    if (!read_file(path, text)) { err = "cannot read " + path; return nullptr; }
    bool have_bin = read_file(bin_path.empty() ? path + "bin" : bin_path, bin);
    JVal j;
    if (!parse_json(text, j, err)) return nullptr;
    auto* s = new luzscene();
    LuzLoad L{s, &bin, have_bin, {}, {}, 0, 0, {}};
    if (const JVal* assets = j.get("assets")) {
        for (const JVal& a : assets->a) {
            int type = (int)(a.get("type") ? a.get("type")->num() : 0);
            uint64_t uuid = a.get("uuid") ? a.get("uuid")->u64() : 0;
            if (type == kMesh) {
                size_t vo, vs, io, is;
                if (!have_bin || !blob_range(a.get("vertices"), bin, vo, vs) || !blob_range(a.get("indices"), bin, io, is))
                    continue;   // mesh data absent: nodes using it are counted as missing
                if (vs % sizeof(MeshVertex) || is % 4) { err = ".luz: mesh sizes not whole vertices/indices"; delete s; return nullptr; }
                MeshAsset m;
                m.vertices.resize(vs / sizeof(MeshVertex));
                m.indices.resize(is / 4);
                if (vs) std::memcpy(m.vertices.data(), bin.data() + vo, vs);
                if (is) std::memcpy(m.indices.data(), bin.data() + io, is);
                if (!valid_mesh(m)) { err = ".luz: mesh index out of range"; delete s; return nullptr; }
                L.mesh_of[uuid] = (uint32_t)s->meshes.size();
                s->meshes.push_back(std::move(m));
            } else if (type == kMaterial) {
                MaterialAsset m;
                if (const JVal* c = a.get("color")) for (size_t k = 0; k < 4 && k < c->a.size(); ++k) m.color[k] = (float)c->a[k].num();
                if (const JVal* c = a.get("emission")) for (size_t k = 0; k < 3 && k < c->a.size(); ++k) m.emission[k] = (float)c->a[k].num();
                if (const JVal* v = a.get("metallic")) m.metallic = (float)v->num();
                if (const JVal* v = a.get("roughness")) m.roughness = (float)v->num();
                L.mat_of[uuid] = (uint32_t)s->materials.size();
                s->materials.push_back(m);
            }
        }
    }
    s->materials.push_back(MaterialAsset());   // for mesh nodes without a material
    const JVal* scenes = j.get("scenes");
    uint64_t initial = j.get("initialScene") ? j.get("initialScene")->u64() : 0;
    const JVal* scene = nullptr;
    if (scenes && scenes->t == JVal::Obj) {
        for (const auto& kv : scenes->o) {
            uint64_t id = kv.second.get("uuid") ? kv.second.get("uuid")->u64() : std::strtoull(kv.first.c_str(), nullptr, 10);
            if (!scene || id == initial) scene = &kv.second;
            if (id == initial) break;
        }
    }
    if (!scene) { err = ".luz: no scene"; delete s; return nullptr; }
    auto f = [&](const char* k, float& v) { if (const JVal* x = scene->get(k)) v = (float)x->num(); };
    auto i = [&](const char* k, int& v) { if (const JVal* x = scene->get(k)) v = (int)x->num(); };
    f("ambientLight", s->ambientLight);
    if (const JVal* c = scene->get("ambientLightColor")) for (size_t k = 0; k < 3 && k < c->a.size(); ++k) s->ambientLightColor[k] = (float)c->a[k].num();
    i("lightSamples", s->lightSamples);
    i("aoSamples", s->aoSamples);
    f("aoMin", s->aoMin);
    f("aoMax", s->aoMax);
    i("shadowType", s->shadowType);
    L.main_camera = scene->get("mainCamera") ? scene->get("mainCamera")->u64() : 0;
    if (const JVal* nodes = scene->get("nodes"))
        for (const JVal& n : nodes->a) L.node(n, M34::identity(), false);
    if (missing) *missing = L.missing_meshes;
    return s;
}

// Writes the scene as one .luz scene (one Node per mesh node whose parent is
// not the identity, then the MeshNode; light nodes; the camera) + .luzbin.
bool save_luz(const luzscene& s, const std::string& path, const std::string& bin_path, std::string& err) {
    std::string bin;
    uint64_t next_uuid = 0x1000;
    auto uuid = [&]() { return next_uuid++; };
    auto push = [&](const void* p, size_t n) { size_t o = bin.size(); bin.append((const char*)p, n); return o; };
    JOut o;
    o.raw("{\"assets\":[");
    std::vector<uint64_t> mesh_uuid, mat_uuid;
    bool first = true;
    for (const MeshAsset& m : s.meshes) {
        uint64_t id = uuid();
        mesh_uuid.push_back(id);
        size_t vo = push(m.vertices.data(), m.vertices.size() * sizeof(MeshVertex));
        size_t io = push(m.indices.data(), m.indices.size() * 4);
        if (!first) o.raw(",");
        first = false;
        o.raw("{\"type\":2,\"name\":\"mesh\",\"uuid\":"); o.u(id);
        o.raw(",\"vertices\":{\"offset\":"); o.u(vo); o.raw(",\"size\":"); o.u(m.vertices.size() * sizeof(MeshVertex));
        o.raw("},\"indices\":{\"offset\":"); o.u(io); o.raw(",\"size\":"); o.u(m.indices.size() * 4); o.raw("}}");
    }
    for (const MaterialAsset& m : s.materials) {
        uint64_t id = uuid();
        mat_uuid.push_back(id);
        if (!first) o.raw(",");
        first = false;
        o.raw("{\"type\":3,\"name\":\"material\",\"uuid\":"); o.u(id);
        o.raw(",\"color\":["); for (int k = 0; k < 4; ++k) { if (k) o.raw(","); o.f(m.color[k]); }
        o.raw("],\"emission\":["); for (int k = 0; k < 3; ++k) { if (k) o.raw(","); o.f(m.emission[k]); }
        o.raw("],\"metallic\":"); o.f(m.metallic); o.raw(",\"roughness\":"); o.f(m.roughness);
        o.raw(",\"colorMap\":0,\"aoMap\":0,\"emissionMap\":0,\"normalMap\":0,\"metallicRoughnessMap\":0}");
    }
    const uint64_t scene_id = uuid(), cam_id = uuid();
    o.raw("],\"initialScene\":"); o.u(scene_id);
    o.raw(",\"scenes\":{\""); o.u(scene_id); o.raw("\":{\"type\":4,\"name\":\"Scene\",\"uuid\":"); o.u(scene_id);
    o.raw(",\"ambientLight\":"); o.f(s.ambientLight);
    o.raw(",\"ambientLightColor\":["); for (int k = 0; k < 3; ++k) { if (k) o.raw(","); o.f(s.ambientLightColor[k]); }
    o.raw("],\"lightSamples\":"); o.i(s.lightSamples); o.raw(",\"aoSamples\":"); o.i(s.aoSamples);
    o.raw(",\"aoMin\":"); o.f(s.aoMin); o.raw(",\"aoMax\":"); o.f(s.aoMax);
    o.raw(",\"shadowType\":"); o.i(s.shadowType); o.raw(",\"mainCamera\":"); o.u(cam_id);
    o.raw(",\"nodes\":[");
    auto trs = [&](const Node& n) {
        o.raw(",\"position\":["); o.d(n.position.x); o.raw(","); o.d(n.position.y); o.raw(","); o.d(n.position.z);
        o.raw("],\"rotation\":["); o.d(n.rotation.x); o.raw(","); o.d(n.rotation.y); o.raw(","); o.d(n.rotation.z);
        o.raw("],\"scale\":["); o.d(n.scale.x); o.raw(","); o.d(n.scale.y); o.raw(","); o.d(n.scale.z); o.raw("]");
    };
    bool fn = true;
    auto open_parent = [&](const Node& n) -> bool {
        if (!n.has_parent) return false;
        // a parent chain is stored as a Node whose TRS reproduces it (no shear)
        const M34& P = n.parent;
        D3 X = d3(P.m[0][0], P.m[1][0], P.m[2][0]), Y = d3(P.m[0][1], P.m[1][1], P.m[2][1]), Z = d3(P.m[0][2], P.m[1][2], P.m[2][2]);
        double sx = length(X), sy = length(Y), sz = length(Z);
        if (dot(cross(X, Y), Z) < 0) sx = -sx;
        double r[3][3] = {{X.x / sx, Y.x / sy, Z.x / sz}, {X.y / sx, Y.y / sy, Z.y / sz}, {X.z / sx, Y.z / sy, Z.z / sz}};
        Quat q;
        double tr = r[0][0] + r[1][1] + r[2][2];
        if (tr > 0) { double k = std::sqrt(tr + 1.0) * 2; q = Quat{0.25 * k, (r[2][1] - r[1][2]) / k, (r[0][2] - r[2][0]) / k, (r[1][0] - r[0][1]) / k}; }
        else if (r[0][0] > r[1][1] && r[0][0] > r[2][2]) { double k = std::sqrt(1.0 + r[0][0] - r[1][1] - r[2][2]) * 2; q = Quat{(r[2][1] - r[1][2]) / k, 0.25 * k, (r[0][1] + r[1][0]) / k, (r[0][2] + r[2][0]) / k}; }
        else if (r[1][1] > r[2][2]) { double k = std::sqrt(1.0 + r[1][1] - r[0][0] - r[2][2]) * 2; q = Quat{(r[0][2] - r[2][0]) / k, (r[0][1] + r[1][0]) / k, 0.25 * k, (r[1][2] + r[2][1]) / k}; }
        else { double k = std::sqrt(1.0 + r[2][2] - r[0][0] - r[1][1]) * 2; q = Quat{(r[1][0] - r[0][1]) / k, (r[0][2] + r[2][0]) / k, (r[1][2] + r[2][1]) / k, 0.25 * k}; }
        Node pn;
        pn.position = d3(P.m[0][3], P.m[1][3], P.m[2][3]);
        pn.rotation = euler_from_quat(q) * (180.0 / kPiD);
        pn.scale = d3(sx, sy, sz);
        if (!fn) o.raw(",");
        fn = false;
        o.raw("{\"type\":5,\"name\":\"Node\",\"uuid\":"); o.u(uuid());
        trs(pn);
        o.raw(",\"children\":[");
        return true;
    };
    for (const MeshNode& n : s.meshNodes) {
        bool wrapped = open_parent(n);
        if (!wrapped && !fn) o.raw(",");
        fn = false;
        o.raw("{\"type\":6,\"name\":\"MeshNode\",\"uuid\":"); o.u(uuid());
        trs(n);
        o.raw(",\"mesh\":"); o.u(n.mesh < mesh_uuid.size() ? mesh_uuid[n.mesh] : 0);
        o.raw(",\"material\":"); o.u(n.material < mat_uuid.size() ? mat_uuid[n.material] : 0);
        o.raw(",\"children\":[]}");
        if (wrapped) o.raw("]}");
    }
    for (const LightNode& L : s.lights) {
        bool wrapped = open_parent(L);
        if (!wrapped && !fn) o.raw(",");
        fn = false;
        o.raw("{\"type\":7,\"name\":\"Light\",\"uuid\":"); o.u(uuid());
        trs(L);
        o.raw(",\"color\":["); for (int k = 0; k < 3; ++k) { if (k) o.raw(","); o.f(L.color[k]); }
        o.raw("],\"intensity\":"); o.f(L.intensity); o.raw(",\"lightType\":"); o.i((int)L.lightType);
        o.raw(",\"innerAngle\":"); o.f(L.innerAngle); o.raw(",\"outerAngle\":"); o.f(L.outerAngle);
        o.raw(",\"radius\":"); o.f(L.radius); o.raw(",\"shadowMapRange\":"); o.f(L.shadowMapRange);
        o.raw(",\"shadowMapFar\":"); o.f(L.shadowMapFar); o.raw(",\"children\":[]}");
        if (wrapped) o.raw("]}");
    }
    if (!fn) o.raw(",");
    const CameraNode& c = s.camera;
    o.raw("{\"type\":8,\"name\":\"Camera\",\"uuid\":"); o.u(cam_id);
    o.raw(",\"mode\":1,\"cameraType\":0,\"eye\":["); o.d(c.eye.x); o.raw(","); o.d(c.eye.y); o.raw(","); o.d(c.eye.z);
    o.raw("],\"center\":["); o.d(c.center.x); o.raw(","); o.d(c.center.y); o.raw(","); o.d(c.center.z);
    o.raw("],\"horizontalFov\":"); o.f(c.horizontalFov); o.raw(",\"nearDistance\":"); o.f(c.nearDistance);
    o.raw(",\"farDistance\":"); o.f(c.farDistance); o.raw(",\"children\":[]}]}}}");
    std::ofstream fj(path, std::ios::binary), fb(bin_path.empty() ? path + "bin" : bin_path, std::ios::binary);
    if (!fj || !fb) { err = "cannot write " + path; return false; }
    fj << o.s;
    fb.write(bin.data(), (std::streamsize)bin.size());
    if (!fj || !fb) { err = "write failed for " + path; return false; }
    return true;
}

void set_error(char* buf, uint32_t len, const std::string& msg) {
    if (!buf || len == 0) return;
    std::snprintf(buf, len, "%s", msg.c_str());
}

}  // namespace
}  // namespace luz

using namespace luz;

extern "C" {

luzscene* luzscene_load(const char* path, const char* bin_path, uint32_t* missing_meshes, char* error,
                        uint32_t error_len) {
    if (missing_meshes) *missing_meshes = 0;
    if (!path) { set_error(error, error_len, "null path"); return nullptr; }
    std::string p(path), err;
    std::string e = ext_of(p);
    luzscene* s = nullptr;
    if (e == ".luz") s = load_luz(p, bin_path ? bin_path : "", err, missing_meshes);
    else if (e == ".gltf" || e == ".glb") s = load_gltf(p, err);
    else if (e == ".obj") s = load_obj(p, err);
    else err = "unknown scene format " + e;
    if (!s) set_error(error, error_len, err);
    return s;
}

int luzscene_save(const luzscene* s, const char* path, const char* bin_path, char* error, uint32_t error_len) {
    if (!s || !path) { set_error(error, error_len, "null argument"); return -1; }
    std::string err;
    if (!save_luz(*s, path, bin_path ? bin_path : "", err)) { set_error(error, error_len, err); return -1; }
    return 0;
}

int luzscene_set_camera(luzscene* s, const float eye[3], const float center[3], float fov_degrees, uint32_t width,
                        uint32_t height, float near_distance, float far_distance) {
    if (!s || !eye || !center || width == 0 || height == 0 || !(fov_degrees > 0.0f && fov_degrees < 180.0f)) return -1;
    if (near_distance < 0.0f || far_distance < 0.0f || (far_distance > 0.0f && far_distance <= near_distance)) return -1;
    if (near_distance > 0.0f) s->camera.nearDistance = near_distance;
    if (far_distance > 0.0f) s->camera.farDistance = far_distance;
    s->camera.eye = d3(eye[0], eye[1], eye[2]);
    s->camera.center = d3(center[0], center[1], center[2]);
    s->camera.horizontalFov = fov_degrees;
    s->width = width;
    s->height = height;
    return 0;
}

int luzscene_add_light(luzscene* s, int type, const float position[3], const float direction[3],
                       const float color[3], float intensity, float radius) {
    if (!s || !position || type < 0 || type > 2 || s->lights.size() >= LUZRT_MAX_LIGHTS) return -1;
    LightNode L;
    L.lightType = (LightNode::LightType)type;
    L.position = d3(position[0], position[1], position[2]);
    if (direction) L.SetDirection(d3(direction[0], direction[1], direction[2]));
    if (color) for (int k = 0; k < 3; ++k) L.color[k] = color[k];
    L.intensity = intensity;
    L.radius = radius;
    s->lights.push_back(L);
    return 0;
}

int luzscene_set_samples(luzscene* s, int light_samples, int ao_samples, float ao_min, float ao_max) {
    if (!s || light_samples < 0 || light_samples > LUZRT_MAX_SAMPLES || ao_samples < 0 || ao_samples > LUZRT_MAX_SAMPLES)
        return -1;
    s->lightSamples = light_samples;
    s->aoSamples = ao_samples;
    s->aoMin = ao_min;
    s->aoMax = ao_max;
    return 0;
}

}  // extern "C"
