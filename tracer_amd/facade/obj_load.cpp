// Wavefront OBJ / MTL ingest behind the reference's load_model_from_file and
// create_texture_from_file (src/utility/mesh_load.cpp:25-113,
// src/utility/image_file.cpp:139-166).
//
// The reference parses with tinyobjloader (fetched at configure time by
// dep/tinyobjloader.cmake; its GIT_TAG v2.0.0rc9 line is commented out, so
// the version is not pinned) and then copies what the reader returns into
// model_t.  This file is its own reader of the same subset of the format,
// following tinyobjloader's documented behaviour:
//   * v / vn / vt records; extra components (w, vertex colours) are ignored;
//   * f records of v, v/vt, v//vn or v/vt/vn corners; indices are 1-based,
//     negative ones count back from the last record read; 0 or out of range
//     is a parse error;
//   * o and g start a new shape once the current one holds faces; a shape is
//     named by the rest of its o / g line;
//   * usemtl sets the material of the faces that follow (-1 when none or
//     unknown) without starting a shape; mtllib loads the named libraries
//     relative to the OBJ file's directory;
//   * faces are triangulated: triangles as given, quads split along their
//     shorter diagonal ((0,1,2),(0,2,3) when |p2-p0| < |p3-p1|, else
//     (0,1,3),(1,2,3)), larger polygons as a fan about their first corner
//     (exact for convex polygons);
//   * MTL: newmtl, Ka, Kd, Ks, Kt/Tf, Ke (a single value sets the first
//     component, the others stay 0), Ns, Ni, d, Tr (d = 1 - Tr), illum, the
//     map_* / bump / disp / refl / norm texture names (the last token of the
//     line), and the PBR extension Pr, Pm, Ps, Pc, Pcr with their maps;
//     defaults as tinyobjloader's InitMaterial (colours 0, Ns 1, Ni 1, d 1).
// Then, as mesh_load.cpp:75-108 does: per shape, corners are deduplicated on
// exact (position, normalised normal, uv) into mesh_t::vertices, with uv 0
// when a corner has none.  A corner without a normal reads past the normal
// array in the reference (mesh_load.cpp:91-93, undefined behaviour); here it
// takes its triangle's geometric normal normalize((p1 - p0) x (p2 - p0)).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "facade_impl.h"

namespace tracer {
namespace {

struct Corner {
    int v = -1, vt = -1, vn = -1;
};

[[noreturn]] void parse_error(const std::string& file, int line, const std::string& what) {
    throw std::runtime_error("load_model_from_file: " + file + ":" + std::to_string(line) + ": " + what);
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

// Splits a line into its keyword and the rest (comments and '\r' removed).
bool split_keyword(std::string line, std::string& key, std::string& rest) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    line = trim(line);
    if (line.empty()) return false;
    size_t i = 0;
    while (i < line.size() && !std::isspace(static_cast<unsigned char>(line[i]))) ++i;
    key = line.substr(0, i);
    rest = trim(line.substr(i));
    return true;
}

// Up to n floats from s; returns how many were read.
int read_floats(const std::string& s, float* out, int n) {
    const char* p = s.c_str();
    int k = 0;
    while (k < n) {
        char* end = nullptr;
        errno = 0;
        const float v = std::strtof(p, &end);
        if (end == p) break;
        out[k++] = v;
        p = end;
    }
    return k;
}

// tinyobjloader's fixIndex: 1-based, negative counts back from the end.
bool fix_index(long idx, size_t count, int& out) {
    if (idx > 0) {
        if (size_t(idx) > count) return false;
        out = int(idx - 1);
        return true;
    }
    if (idx < 0) {
        if (size_t(-idx) > count) return false;
        out = int(long(count) + idx);
        return true;
    }
    return false;
}

bool parse_corner(const std::string& tok, size_t nv, size_t nvt, size_t nvn, Corner& c) {
    const char* p = tok.c_str();
    char* end = nullptr;
    const long v = std::strtol(p, &end, 10);
    if (end == p || !fix_index(v, nv, c.v)) return false;
    p = end;
    if (*p == '\0') return true;
    if (*p != '/') return false;
    ++p;
    if (*p != '/') {
        const long t = std::strtol(p, &end, 10);
        if (end == p || !fix_index(t, nvt, c.vt)) return false;
        p = end;
    }
    if (*p == '\0') return true;
    if (*p != '/') return false;
    ++p;
    const long n = std::strtol(p, &end, 10);
    if (end == p || !fix_index(n, nvn, c.vn)) return false;
    return *end == '\0';
}

std::string dir_of(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

// mesh_load.cpp:8-23 (keeps the separator, as the reference does)
std::string extract_name_from_path(const std::string& path) {
    const size_t s = path.find_last_of("/\\");
    return s == std::string::npos ? path : path.substr(s);
}

std::string texture_name(const std::string& rest) {
    // texture options (-bm 1, -o u v w, ...) come first; the file name is the last token
    std::istringstream in(rest);
    std::string tok, last;
    while (in >> tok) last = tok;
    return last;
}

void load_mtl(const std::string& path, std::vector<material_t>& mats, std::unordered_map<std::string, int>& ids) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("load_model_from_file: cannot open material library " + path);
    std::string line, key, rest;
    material_t* cur = nullptr;
    int ln = 0;
    auto rgb = [&](real* dst) {
        float v[3] = {0, 0, 0};
        if (read_floats(rest, v, 3) < 1) parse_error(path, ln, "expected a colour");
        for (int k = 0; k < 3; ++k) dst[k] = v[k];
    };
    // a texture found next to the MTL file is read from there, else the name is
    // used as given (the reference opens it relative to the working directory)
    auto tn = [&]() {
        const std::string n = texture_name(rest);
        if (!n.empty() && n[0] != '/') {
            const std::string j = dir_of(path) + n;
            if (std::ifstream(j).good()) return j;
        }
        return n;
    };
    auto scalar = [&]() {
        float v = 0;
        if (read_floats(rest, &v, 1) != 1) parse_error(path, ln, "expected a number");
        return v;
    };
    while (std::getline(in, line)) {
        ++ln;
        if (!split_keyword(line, key, rest)) continue;
        if (key == "newmtl") {
            if (ids.count(rest)) {   // a second definition of a name is ignored
                cur = nullptr;
                continue;
            }
            ids[rest] = int(mats.size());
            mats.emplace_back();
            cur = &mats.back();
            cur->name = rest;
            continue;
        }
        if (!cur) continue;
        if (key == "Ka") rgb(cur->ambient);
        else if (key == "Kd") rgb(cur->diffuse);
        else if (key == "Ks") rgb(cur->specular);
        else if (key == "Kt" || key == "Tf") rgb(cur->transmittance);
        else if (key == "Ke") rgb(cur->emission);
        else if (key == "Ns") cur->shininess = scalar();
        else if (key == "Ni") cur->ior = scalar();
        else if (key == "d") cur->dissolve = scalar();
        else if (key == "Tr") cur->dissolve = 1.0f - scalar();
        else if (key == "illum") cur->illum = int(scalar());
        else if (key == "Pr") cur->roughness = scalar();
        else if (key == "Pm") cur->metallic = scalar();
        else if (key == "Ps") cur->sheen = scalar();
        else if (key == "Pc") cur->clearcoat_thickness = scalar();
        else if (key == "Pcr") cur->clearcoat_roughness = scalar();
        else if (key == "map_Ka") cur->map_ka = tn();
        else if (key == "map_Kd") cur->map_kd = tn();
        else if (key == "map_Ks") cur->map_ks = tn();
        else if (key == "map_Ns") cur->map_ns = tn();
        else if (key == "map_bump" || key == "map_Bump" || key == "bump") cur->map_bump = tn();
        else if (key == "disp") cur->disp = tn();
        else if (key == "map_d") cur->map_d = tn();
        else if (key == "refl") cur->refl = tn();
        else if (key == "map_Pr") cur->map_pr = tn();
        else if (key == "map_Pm") cur->map_pm = tn();
        else if (key == "map_Ps") cur->map_ps = tn();
        else if (key == "map_Ke") cur->map_ke = tn();
        else if (key == "norm") cur->norm = tn();
    }
}

struct Face {
    Corner c[3];
    int material;
};

struct ShapeFaces {
    std::string name;
    std::vector<Face> tris;
};

struct VertexKey {
    float v[8];
    bool operator==(const VertexKey& o) const { return std::memcmp(v, o.v, sizeof(v)) == 0; }
};
struct VertexKeyHash {
    size_t operator()(const VertexKey& k) const {
        uint64_t h = 1469598103934665603ull;   // FNV-1a over the bytes
        const unsigned char* b = reinterpret_cast<const unsigned char*>(k.v);
        for (size_t i = 0; i < sizeof(k.v); ++i) h = (h ^ b[i]) * 1099511628211ull;
        return size_t(h);
    }
};

}  // namespace

model_t load_model_from_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("load_model_from_file: cannot open " + path);
    model_t model;
    model.name = extract_name_from_path(path);
    std::vector<float> P, N, T;
    std::unordered_map<std::string, int> mat_ids;
    std::vector<ShapeFaces> shapes(1);
    int material = -1;
    std::string line, key, rest;
    int ln = 0;
    std::vector<Corner> poly;
    while (std::getline(in, line)) {
        ++ln;
        if (!split_keyword(line, key, rest)) continue;
        if (key == "v") {
            float v[3];
            if (read_floats(rest, v, 3) != 3) parse_error(path, ln, "vertex needs 3 coordinates");
            P.insert(P.end(), v, v + 3);
        } else if (key == "vn") {
            float v[3];
            if (read_floats(rest, v, 3) != 3) parse_error(path, ln, "normal needs 3 components");
            N.insert(N.end(), v, v + 3);
        } else if (key == "vt") {
            float v[2] = {0, 0};
            if (read_floats(rest, v, 2) < 1) parse_error(path, ln, "texture coordinate needs a component");
            T.insert(T.end(), v, v + 2);
        } else if (key == "f") {
            poly.clear();
            std::istringstream toks(rest);
            std::string tok;
            while (toks >> tok) {
                Corner c;
                if (!parse_corner(tok, P.size() / 3, T.size() / 2, N.size() / 3, c))
                    parse_error(path, ln, "invalid face corner '" + tok + "'");
                poly.push_back(c);
            }
            if (poly.size() < 3) continue;   // points and lines carry no triangle
            auto& tris = shapes.back().tris;
            if (poly.size() == 4) {
                auto pos = [&](int i) { const float* p = &P[size_t(poly[size_t(i)].v) * 3]; return Vector3f(p[0], p[1], p[2]); };
                const Vector3f e02 = pos(2) - pos(0), e13 = pos(3) - pos(1);
                if (dot(e02, e02) < dot(e13, e13)) {
                    tris.push_back({{poly[0], poly[1], poly[2]}, material});
                    tris.push_back({{poly[0], poly[2], poly[3]}, material});
                } else {
                    tris.push_back({{poly[0], poly[1], poly[3]}, material});
                    tris.push_back({{poly[1], poly[2], poly[3]}, material});
                }
            } else {
                for (size_t k = 2; k < poly.size(); ++k) tris.push_back({{poly[0], poly[k - 1], poly[k]}, material});
            }
        } else if (key == "o" || key == "g") {
            if (!shapes.back().tris.empty()) shapes.emplace_back();
            shapes.back().name = rest;
        } else if (key == "usemtl") {
            const auto it = mat_ids.find(rest);
            material = it == mat_ids.end() ? -1 : it->second;
        } else if (key == "mtllib") {
            std::istringstream libs(rest);
            std::string lib;
            while (libs >> lib) load_mtl(dir_of(path) + lib, model.material, mat_ids);
        }
    }
    for (const ShapeFaces& sf : shapes) {
        if (sf.tris.empty()) continue;
        model.mesh.emplace_back();
        mesh_t& mesh = model.mesh.back();
        mesh.name = sf.name;
        std::unordered_map<VertexKey, int, VertexKeyHash> unique;
        for (const Face& f : sf.tris) {
            Vector3f p[3];
            for (int k = 0; k < 3; ++k) {
                const float* q = &P[size_t(f.c[k].v) * 3];
                p[k] = Vector3f(q[0], q[1], q[2]);
            }
            const Vector3f face_n = cross(p[1] - p[0], p[2] - p[0]).normalize();
            for (int k = 0; k < 3; ++k) {
                vertex_t vx;
                vx.pos = Point3f(p[k].x, p[k].y, p[k].z);
                if (f.c[k].vn >= 0) {
                    const float* q = &N[size_t(f.c[k].vn) * 3];
                    vx.n = Normal3f(q[0], q[1], q[2]).normalize();
                } else {
                    vx.n = Normal3f(face_n.x, face_n.y, face_n.z);
                }
                if (f.c[k].vt >= 0) vx.uv = Point2f(T[size_t(f.c[k].vt) * 2], T[size_t(f.c[k].vt) * 2 + 1]);
                // vertex_t::operator== compares floats, so -0 and +0 are one value: x + 0.0f maps -0 to +0
                const VertexKey key{{vx.pos.x + 0.0f, vx.pos.y + 0.0f, vx.pos.z + 0.0f, vx.n.x + 0.0f, vx.n.y + 0.0f,
                                     vx.n.z + 0.0f, vx.uv.x + 0.0f, vx.uv.y + 0.0f}};
                const auto it = unique.find(key);
                if (it != unique.end()) {
                    mesh.indices.push_back(it->second);
                } else {
                    const int id = int(mesh.vertices.size());
                    unique.emplace(key, id);
                    mesh.vertices.push_back(vx);
                    mesh.indices.push_back(id);
                }
            }
            mesh.materials.push_back(f.material);
        }
    }
    return model;
}

namespace {
RC<Texture2D> constant_or_file(const std::string& name, const real* rgb) {
    if (!name.empty()) return detail::texture_from_file_scaled(name, rgb);
    return create_constant_texture2d(Spectrum(rgb[0], rgb[1], rgb[2]));
}
}  // namespace

MaterialTexture create_texture_from_file(const material_t& m) {
    MaterialTexture t;
    const real ns[3] = {m.shininess, m.shininess, m.shininess};
    t.map_ka = constant_or_file(m.map_ka, m.ambient);
    t.map_kd = constant_or_file(m.map_kd, m.diffuse);
    t.map_ks = constant_or_file(m.map_ks, m.specular);
    t.map_ns = constant_or_file(m.map_ns, ns);
    t.map_ke = constant_or_file(m.map_ke, m.emission);
    if (m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0) t.has_emission = true;
    if (m.ior > 1) {
        if (t.map_ks->evaluate_s(Point2f(0.5f, 0.5f)) == 0) t.map_ks = create_constant_texture2d(Spectrum(1.0f));
        t.map_kt = t.map_ks;
        t.as_glass = true;
        t.ior = create_constant_texture2d(Spectrum(m.ior));
    }
    return t;
}

}  // namespace tracer
