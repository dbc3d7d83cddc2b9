// Synthetic benchmark scenes (configurations 1-5), assembled through the
// tracer:: factory API the way the reference's src/main.cpp assembles its
// scenes, then flattened for the C ABI.
//
// Draw order (PCG32, pcg32_srandom(seed, 0x5eed)):
//   c1: yaw of inner box A, yaw of inner box B, then albedo r,g,b per face in
//       the order room faces (floor, ceiling, back, left, right), box A faces,
//       box B faces, light quad.
//   sphere set (c2, c3, c4, c5): per sphere: centre x,y,z, scale, noise
//       offset x,y,z, base colour r,g,b, metallic draw, roughness, clearcoat,
//       sheen, specTrans draw; then per quad light: x, z, y, radiance.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>

#include "facade_impl.h"
#include "tracer/scenes.h"

namespace tracer {
namespace scenes {
namespace {

struct Rng {
    uint64_t state = 0, inc = 1;
    explicit Rng(uint64_t seed) {
        inc = (0x5eedULL << 1) | 1u;
        state = 0;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    float uni() { return float(next() >> 8) * (1.0f / 16777216.0f); }
    float uni(float a, float b) { return a + (b - a) * uni(); }
};

// 3D simplex noise (Gustavson, "Simplex noise demystified", 2005), with a
// permutation drawn from the scene RNG.
class Simplex {
public:
    explicit Simplex(Rng& r) {
        for (int i = 0; i < 256; ++i) p[i] = uint8_t(i);
        for (int i = 255; i > 0; --i) std::swap(p[i], p[r.next() % uint32_t(i + 1)]);
        for (int i = 0; i < 512; ++i) perm[i] = p[i & 255];
    }
    float noise(float x, float y, float z) const {
        static const int g[12][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
                                     {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
        const float F3 = 1.0f / 3.0f, G3 = 1.0f / 6.0f;
        const float s = (x + y + z) * F3;
        const int i = int(std::floor(x + s)), j = int(std::floor(y + s)), k = int(std::floor(z + s));
        const float t = (i + j + k) * G3;
        const float x0 = x - (i - t), y0 = y - (j - t), z0 = z - (k - t);
        int i1, j1, k1, i2, j2, k2;
        if (x0 >= y0) {
            if (y0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
            else if (x0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1; }
            else { i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1; }
        } else {
            if (y0 < z0) { i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1; }
            else if (x0 < z0) { i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1; }
            else { i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
        }
        const float xs[4] = {x0, x0 - i1 + G3, x0 - i2 + 2 * G3, x0 - 1 + 3 * G3};
        const float ys[4] = {y0, y0 - j1 + G3, y0 - j2 + 2 * G3, y0 - 1 + 3 * G3};
        const float zs[4] = {z0, z0 - k1 + G3, z0 - k2 + 2 * G3, z0 - 1 + 3 * G3};
        const int ii = i & 255, jj = j & 255, kk = k & 255;
        const int gi[4] = {perm[ii + perm[jj + perm[kk]]] % 12, perm[ii + i1 + perm[jj + j1 + perm[kk + k1]]] % 12,
                           perm[ii + i2 + perm[jj + j2 + perm[kk + k2]]] % 12, perm[ii + 1 + perm[jj + 1 + perm[kk + 1]]] % 12};
        float n = 0;
        for (int c = 0; c < 4; ++c) {
            float tt = 0.6f - xs[c] * xs[c] - ys[c] * ys[c] - zs[c] * zs[c];
            if (tt > 0) {
                tt *= tt;
                n += tt * tt * (g[gi[c]][0] * xs[c] + g[gi[c]][1] * ys[c] + g[gi[c]][2] * zs[c]);
            }
        }
        return 32.0f * n;
    }
private:
    uint8_t p[256];
    uint8_t perm[512];
};

RC<Material> lambert(const Spectrum& albedo) {
    return create_phong_material(create_constant_texture2d(Spectrum(0)), create_constant_texture2d(albedo),
                                 create_constant_texture2d(Spectrum(0)), create_constant_texture2d(Spectrum(1)));
}

struct DisneyDraw {
    Spectrum base;
    float metallic, roughness, clearcoat, sheen, spec_trans;
};

// Config-2 Disney mapping: specular 0.5 is the Disney "0.08 * specular"
// reflectance 0.04, which the physically based dielectric Fresnel of this
// BSDF already gives at IOR 1.5, so specular_scale = 1; specTrans -> transmission
// with transmission_roughness = roughness; anisotropic = specular_tint = 0,
// sheen_tint = 0.5, clearcoat_gloss = 1.
RC<Material> disney(const DisneyDraw& d) {
    auto c = [](float v) { return create_constant_texture2d(Spectrum(v)); };
    return create_disney(create_constant_texture2d(d.base), c(d.metallic), c(d.roughness), c(d.spec_trans), c(d.roughness),
                         c(1.5f), c(1.0f), c(0.0f), c(0.0f), c(d.sheen), c(0.5f), c(d.clearcoat), c(1.0f),
                         newBox<const NormalMapper>(nullptr), nullptr);
}

struct Builder {
    std::vector<RC<Primitive>> prims;
    std::vector<const Light*> lights;
    std::vector<RC<Light>> owned;
    MediumInterface mi{create_vacuum(), create_vacuum()};

    void add_mesh(const mesh_t& mesh, const Transform& t, const std::vector<RC<Material>>& mat_per_tri,
                  const Spectrum& emission = Spectrum(0)) {
        auto shapes = create_triangle_mesh(mesh, t);
        for (size_t i = 0; i < shapes.size(); ++i) {
            auto p = create_geometric_primitive(shapes[i], mat_per_tri.size() == 1 ? mat_per_tri[0] : mat_per_tri[i], mi, emission);
            if (auto l = p->as_area_light()) lights.push_back(l);
            prims.push_back(std::move(p));
        }
    }
};

// A quad: centre c, edges u and v; the geometric normal is cross(u, v).
void add_quad(mesh_t& m, Vector3f c, Vector3f u, Vector3f v) {
    const Vector3f n = cross(u, v).normalize();
    const int b = int(m.vertices.size());
    const Vector3f cs[4] = {c - u * 0.5f - v * 0.5f, c + u * 0.5f - v * 0.5f, c + u * 0.5f + v * 0.5f, c - u * 0.5f + v * 0.5f};
    const Point2f uvs[4] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    for (int k = 0; k < 4; ++k) m.vertices.push_back({Point3f(cs[k].x, cs[k].y, cs[k].z), Normal3f(n.x, n.y, n.z), uvs[k]});
    m.indices.insert(m.indices.end(), {b, b + 1, b + 2, b, b + 2, b + 3});
}

// Axis-aligned box faces [floor(-y), ceiling(+y), back(-z), left(-x), right(+x), front(+z)],
// normals outward (inward = false) or inward.
mesh_t box_faces(Vector3f lo, Vector3f hi, bool inward, int nfaces) {
    mesh_t m;
    const Vector3f c = (lo + hi) * 0.5f, e = hi - lo;
    const float s = inward ? -1.0f : 1.0f;
    const Vector3f X(e.x, 0, 0), Y(0, e.y, 0), Z(0, 0, e.z);
    // each (u, v) pair is ordered so that cross(u, v) points outward; inward flips v
    struct F { Vector3f c, u, v; };
    const F faces[6] = {
        {Vector3f(c.x, lo.y, c.z), X, Z},        // floor: cross(X, Z) = -y  (outward)
        {Vector3f(c.x, hi.y, c.z), Z, X},        // ceiling: +y
        {Vector3f(c.x, c.y, lo.z), Y, X},        // back: -z
        {Vector3f(lo.x, c.y, c.z), Z, Y},        // left: -x
        {Vector3f(hi.x, c.y, c.z), Y, Z},        // right: +x
        {Vector3f(c.x, c.y, hi.z), X, Y},        // front: +z
    };
    for (int f = 0; f < nfaces; ++f) add_quad(m, faces[f].c, faces[f].u, faces[f].v * s);
    return m;
}

mesh_t transform_mesh(const mesh_t& m, const Transform& t) {
    mesh_t r = m;
    for (auto& v : r.vertices) {
        v.pos = t(v.pos);
        v.n = t(v.n).normalize();
    }
    return r;
}

// Icosphere of the given subdivision level (level 5: 10,242 vertices,
// 20,480 triangles), outward winding.
void icosphere(int level, std::vector<Vector3f>& V, std::vector<std::array<int, 3>>& F) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    V = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& v : V) v = v.normalize();
    F = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
         {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V.push_back(((V[a] + V[b]) * 0.5f).normalize());
            const int id = int(V.size()) - 1;
            mid[key] = id;
            return id;
        };
        std::vector<std::array<int, 3>> nf;
        nf.reserve(F.size() * 4);
        for (auto& f : F) {
            const int ab = midpoint(f[0], f[1]), bc = midpoint(f[1], f[2]), ca = midpoint(f[2], f[0]);
            nf.push_back({f[0], ab, ca});
            nf.push_back({f[1], bc, ab});
            nf.push_back({f[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        F.swap(nf);
    }
    for (auto& f : F) {
        const Vector3f n = cross(V[f[1]] - V[f[0]], V[f[2]] - V[f[0]]);
        if (dot(n, V[f[0]] + V[f[1]] + V[f[2]]) < 0) std::swap(f[1], f[2]);
    }
}

// Unit icosphere with radial displacement 1 + 0.1 * simplex(p + offset) and
// area-weighted vertex normals.
mesh_t displaced_sphere(const std::vector<Vector3f>& V, const std::vector<std::array<int, 3>>& F, const Simplex& noise,
                        Vector3f off) {
    mesh_t m;
    m.vertices.resize(V.size());
    for (size_t i = 0; i < V.size(); ++i) {
        const Vector3f p = V[i];
        const float r = 1.0f + 0.1f * noise.noise(1.5f * p.x + off.x, 1.5f * p.y + off.y, 1.5f * p.z + off.z);
        m.vertices[i].pos = Point3f(p.x * r, p.y * r, p.z * r);
        const float u = std::atan2(p.z, p.x) / (2 * PI_r) + 0.5f;
        const float v = std::acos(std::max(-1.0f, std::min(1.0f, p.y))) / PI_r;
        m.vertices[i].uv = Point2f(u, v);
    }
    std::vector<Vector3f> acc(V.size());
    m.indices.reserve(F.size() * 3);
    for (auto& f : F) {
        const Point3f& a = m.vertices[f[0]].pos;
        const Point3f& b = m.vertices[f[1]].pos;
        const Point3f& c = m.vertices[f[2]].pos;
        const Vector3f n = cross(b - a, c - a);
        for (int k = 0; k < 3; ++k) acc[f[k]] = acc[f[k]] + n;
        m.indices.insert(m.indices.end(), {f[0], f[1], f[2]});
    }
    for (size_t i = 0; i < V.size(); ++i) {
        const Vector3f n = acc[i].normalize();
        m.vertices[i].n = Normal3f(n.x, n.y, n.z);
    }
    return m;
}

struct SphereDraw {
    Vector3f center;
    float scale;
    Vector3f offset;
    DisneyDraw mat;
};

std::vector<SphereDraw> draw_spheres(Rng& r, int count, float extent, float smin, float smax) {
    std::vector<SphereDraw> out(static_cast<size_t>(count));
    for (auto& s : out) {
        s.center = Vector3f(r.uni(-extent, extent), r.uni(-extent, extent), r.uni(-extent, extent));
        s.scale = r.uni(smin, smax);
        s.offset = Vector3f(r.uni(0, 100), r.uni(0, 100), r.uni(0, 100));
        s.mat.base = Spectrum(r.uni(), r.uni(), r.uni());
        s.mat.metallic = r.uni() < 0.3f ? 1.0f : 0.0f;
        s.mat.roughness = r.uni(0.05f, 1.0f);
        s.mat.clearcoat = r.uni();
        s.mat.sheen = r.uni(0.0f, 0.5f);
        s.mat.spec_trans = r.uni() < 0.1f ? 1.0f : 0.0f;
    }
    return out;
}

void add_spheres(Builder& b, const std::vector<SphereDraw>& S, const Simplex& noise, float center_scale) {
    std::vector<Vector3f> V;
    std::vector<std::array<int, 3>> F;
    icosphere(5, V, F);
    for (const auto& s : S) {
        mesh_t m = displaced_sphere(V, F, noise, s.offset);
        const Transform t = translate(s.center * center_scale) * scale(s.scale, s.scale, s.scale);
        b.add_mesh(m, t, {disney(s.mat)});
    }
}

}  // namespace

struct Config {
    RC<Scene> scene;
    Builder builder;
    int width = 256, height = 256, spp = 16, min_depth = 3, max_depth = 8;
    RC<Filter> filter;
    int max_leaf = 3;
};

static void finish(Config& c, RC<Camera> cam, RC<EnvironmentLight> env) {
    auto accel = create_bvh_accel(c.max_leaf);
    accel->build(c.builder.prims);
    c.scene = create_general_scene(accel);
    c.scene->set_camera(cam);
    c.scene->lights = c.builder.lights;
    c.scene->environment_light = env;
    c.scene->prepare_to_render();
}

// c1: tiny procedural room (10 + 24 + 2 triangles).  The inner boxes float
// 2e-3 above the floor so that no two faces are coplanar.
static void config1(Config& c, uint64_t seed) {
    Rng r(seed);
    const float yawA = r.uni(0.0f, PI_r / 4), yawB = r.uni(0.0f, PI_r / 4);
    Builder& b = c.builder;
    std::vector<RC<Material>> room_m, a_m, b_m;
    auto face_mat = [&]() { const float x = r.uni(0.1f, 0.9f), y = r.uni(0.1f, 0.9f), z = r.uni(0.1f, 0.9f); return lambert(Spectrum(x, y, z)); };
    mesh_t room = box_faces(Vector3f(-1, -1, -1), Vector3f(1, 1, 1), true, 5);   // +z face open
    for (int f = 0; f < 5; ++f) { auto m = face_mat(); room_m.push_back(m); room_m.push_back(m); }
    mesh_t ba = box_faces(Vector3f(-0.3f, -0.6f, -0.3f), Vector3f(0.3f, 0.6f, 0.3f), false, 6);
    for (int f = 0; f < 6; ++f) { auto m = face_mat(); a_m.push_back(m); a_m.push_back(m); }
    mesh_t bb = box_faces(Vector3f(-0.3f, -0.3f, -0.3f), Vector3f(0.3f, 0.3f, 0.3f), false, 6);
    for (int f = 0; f < 6; ++f) { auto m = face_mat(); b_m.push_back(m); b_m.push_back(m); }
    auto lm = face_mat();
    b.add_mesh(room, Transform(), room_m);
    b.add_mesh(ba, translate(Vector3f(-0.4f, -0.4f + 2e-3f, -0.3f)) * rotate_y(yawA), a_m);
    b.add_mesh(bb, translate(Vector3f(0.4f, -0.7f + 2e-3f, 0.25f)) * rotate_y(yawB), b_m);
    mesh_t light;
    add_quad(light, Vector3f(0, 1 - 1e-3f, 0), Vector3f(0.5f, 0, 0), Vector3f(0, 0, 0.5f));   // normal -y
    b.add_mesh(light, Transform(), {lm}, Spectrum(17.0f));
    c.width = c.height = 256;
    c.spp = 16;
    c.max_depth = 8;
    c.min_depth = 3;
    c.filter = create_box_filter(0.5f);
    auto cam = create_thin_lens_camera(1.0f, Point3f(0, 0, 3.75f), Point3f(0, 0, 0), Vector3f(0, 1, 0), 40.0f * PI_r / 180.0f, 0.0f, 1.0f);
    finish(c, cam, nullptr);
}

static void add_ceiling_lights(Builder& b, Rng& r, int n, float extent, float ylo, float yhi, float size, float Llo, float Lhi) {
    auto black = lambert(Spectrum(0.0f));
    for (int i = 0; i < n; ++i) {
        const float x = r.uni(-extent, extent), z = r.uni(-extent, extent), y = r.uni(ylo, yhi);
        const float L = r.uni(Llo, Lhi);
        mesh_t q;
        add_quad(q, Vector3f(x, y, z), Vector3f(size, 0, 0), Vector3f(0, 0, size));   // faces -y
        b.add_mesh(q, Transform(), {black}, Spectrum(L));
    }
}

// c2: 64 displaced level-5 icospheres, ground plane, 16 quad lights, constant environment.
static void config2(Config& c, uint64_t seed, double detail) {
    Rng r(seed);
    Simplex noise(r);
    const int count = std::max(1, int(std::lround(64 * detail)));
    auto S = draw_spheres(r, count, 10.0f, 0.2f, 1.0f);
    add_spheres(c.builder, S, noise, 1.0f);
    mesh_t ground;
    add_quad(ground, Vector3f(0, -11, 0), Vector3f(0, 0, 200), Vector3f(200, 0, 0));   // +y
    c.builder.add_mesh(ground, Transform(), {lambert(Spectrum(0.5f))});
    add_ceiling_lights(c.builder, r, 16, 10.0f, 11.0f, 13.0f, 1.5f, 5.0f, 20.0f);
    c.width = c.height = 1024;
    c.spp = 64;
    c.max_depth = 16;
    c.min_depth = 3;
    c.filter = create_box_filter(0.5f);
    auto cam = create_thin_lens_camera(1.0f, Point3f(0, 6, 34), Point3f(0, -2, 0), Vector3f(0, 1, 0), 45.0f * PI_r / 180.0f, 0.0f, 1.0f);
    finish(c, cam, create_ibl_light(create_constant_texture2d(Spectrum(0.5f)), Transform()));
}

// c3: the c1 room closed and scaled x10 with 12 c2 spheres, filled with a
// homogeneous medium (sigma_a 0.05, sigma_s 0.2, g 0.7).
static void config3(Config& c, uint64_t seed, double detail) {
    Rng r0(0);   // the c1 draws
    const float yawA = r0.uni(0.0f, PI_r / 4), yawB = r0.uni(0.0f, PI_r / 4);
    auto fog = create_homogeneous_medium(Spectrum(0.05f), Spectrum(0.2f), 0.7f, 0x7fffffff);
    c.builder.mi = MediumInterface{create_vacuum(), fog};
    std::vector<RC<Material>> room_m, a_m, b_m;
    auto face_mat = [&]() { const float x = r0.uni(0.1f, 0.9f), y = r0.uni(0.1f, 0.9f), z = r0.uni(0.1f, 0.9f); return lambert(Spectrum(x, y, z)); };
    for (int f = 0; f < 5; ++f) { auto m = face_mat(); room_m.push_back(m); room_m.push_back(m); }
    for (int f = 0; f < 6; ++f) { auto m = face_mat(); a_m.push_back(m); a_m.push_back(m); }
    for (int f = 0; f < 6; ++f) { auto m = face_mat(); b_m.push_back(m); b_m.push_back(m); }
    auto lm = face_mat();
    room_m.push_back(room_m[0]);   // front wall closes the room
    room_m.push_back(room_m[0]);
    const float S = 10.0f;
    c.builder.add_mesh(box_faces(Vector3f(-S, -S, -S), Vector3f(S, S, S), true, 6), Transform(), room_m);
    c.builder.add_mesh(box_faces(Vector3f(-3, -6, -3), Vector3f(3, 6, 3), false, 6), translate(Vector3f(-4, -4 + 0.02f, -3)) * rotate_y(yawA), a_m);
    c.builder.add_mesh(box_faces(Vector3f(-3, -3, -3), Vector3f(3, 3, 3), false, 6), translate(Vector3f(4, -7 + 0.02f, 2.5f)) * rotate_y(yawB), b_m);
    mesh_t light;
    add_quad(light, Vector3f(0, S - 1e-2f, 0), Vector3f(5, 0, 0), Vector3f(0, 0, 5));   // faces -y
    c.builder.add_mesh(light, Transform(), {lm}, Spectrum(17.0f));
    Rng r(seed);
    Simplex noise(r);
    auto all = draw_spheres(r, 64, 10.0f, 0.2f, 1.0f);
    const int count = std::max(1, int(std::lround(12 * std::min(1.0, detail * 1.0))));
    all.resize(size_t(count));
    add_spheres(c.builder, all, noise, 0.8f);
    c.width = c.height = 1024;
    c.spp = 128;
    c.max_depth = 32;
    c.min_depth = 3;
    c.filter = create_box_filter(0.5f);
    auto cam = create_thin_lens_camera(1.0f, Point3f(0, 0, 9), Point3f(0, -1, 0), Vector3f(0, 1, 0), 55.0f * PI_r / 180.0f, 0.0f, 1.0f);
    finish(c, cam, nullptr);
}

// c4: the c2 spheres in a closed 30^3 room lit only through a 0.5 x 0.5
// aperture in an inner wall by a 4 x 4 emitter (radiance 200).
static void config4(Config& c, uint64_t seed, double detail) {
    Rng r(1);   // the c2 sphere set
    Simplex noise(r);
    const int count = std::max(1, int(std::lround(64 * detail)));
    auto S = draw_spheres(r, count, 10.0f, 0.2f, 1.0f);
    (void)seed;
    add_spheres(c.builder, S, noise, 1.0f);
    auto wall = lambert(Spectrum(0.7f));
    const float R = 15.0f, W = 13.0f, h = 0.25f;
    c.builder.add_mesh(box_faces(Vector3f(-R, -R, -R), Vector3f(R, R, R), true, 6), Transform(), {wall});
    mesh_t ground;
    add_quad(ground, Vector3f(-1, -11, 0), Vector3f(0, 0, 2 * R), Vector3f(W + R, 0, 0));
    c.builder.add_mesh(ground, Transform(), {lambert(Spectrum(0.5f))});
    // inner wall at x = W facing -x, with a square hole of half-size h at y = z = 0
    mesh_t inner;
    auto rect = [&](float y0, float y1, float z0, float z1) {
        add_quad(inner, Vector3f(W, 0.5f * (y0 + y1), 0.5f * (z0 + z1)), Vector3f(0, 0, z1 - z0), Vector3f(0, y1 - y0, 0));
    };
    rect(-R, -h, -R, R);
    rect(h, R, -R, R);
    rect(-h, h, -R, -h);
    rect(-h, h, h, R);
    c.builder.add_mesh(inner, Transform(), {wall});
    mesh_t light;
    add_quad(light, Vector3f(14.5f, 0, 0), Vector3f(0, 0, 4), Vector3f(0, 4, 0));   // faces -x
    c.builder.add_mesh(light, Transform(), {lambert(Spectrum(0.0f))}, Spectrum(200.0f));
    c.width = c.height = 1024;
    c.spp = 256;
    c.max_depth = 16;
    c.min_depth = 3;
    c.filter = create_box_filter(0.5f);
    auto cam = create_thin_lens_camera(1.0f, Point3f(-12, 4, 12), Point3f(0, -2, 0), Vector3f(0, 1, 0), 60.0f * PI_r / 180.0f, 0.0f, 1.0f);
    finish(c, cam, nullptr);
}

// c5: 512 displaced level-5 icospheres over a 2048^2-quad fBm terrain, thin
// fog (sigma_s 0.02, g 0.3), 256 quad lights and a sun.
static void config5(Config& c, uint64_t seed, double detail) {
    Rng r(seed);
    Simplex noise(r);
    auto fog = create_homogeneous_medium(Spectrum(0.0f), Spectrum(0.02f), 0.3f, 0x7fffffff);
    c.builder.mi = MediumInterface{create_vacuum(), fog};
    const int grid = std::max(8, int(std::lround(2048 * std::sqrt(detail))));
    const float E = 100.0f;
    auto height = [&](float x, float z) {
        float hgt = 0, amp = 8, f = 0.02f;
        for (int o = 0; o < 6; ++o) { hgt += amp * noise.noise(x * f, 0.37f * o, z * f); amp *= 0.5f; f *= 2.0f; }
        return hgt;
    };
    mesh_t terrain;
    terrain.vertices.resize(size_t(grid + 1) * (grid + 1));
    for (int j = 0; j <= grid; ++j)
        for (int i = 0; i <= grid; ++i) {
            const float x = -E + 2 * E * i / grid, z = -E + 2 * E * j / grid;
            auto& v = terrain.vertices[size_t(j) * (grid + 1) + i];
            v.pos = Point3f(x, height(x, z), z);
            v.uv = Point2f(float(i) / grid, float(j) / grid);
        }
    const float dx = 2 * E / grid;
    for (int j = 0; j <= grid; ++j)
        for (int i = 0; i <= grid; ++i) {
            auto& v = terrain.vertices[size_t(j) * (grid + 1) + i];
            const float hx = height(v.pos.x + dx, v.pos.z) - height(v.pos.x - dx, v.pos.z);
            const float hz = height(v.pos.x, v.pos.z + dx) - height(v.pos.x, v.pos.z - dx);
            const Vector3f n = Vector3f(-hx, 2 * dx, -hz).normalize();
            v.n = Normal3f(n.x, n.y, n.z);
        }
    terrain.indices.reserve(size_t(grid) * grid * 6);
    for (int j = 0; j < grid; ++j)
        for (int i = 0; i < grid; ++i) {
            const int a = j * (grid + 1) + i, b = a + 1, cc = a + grid + 1, d = cc + 1;
            terrain.indices.insert(terrain.indices.end(), {a, cc, b, b, cc, d});   // normals +y
        }
    c.builder.add_mesh(terrain, Transform(), {lambert(Spectrum(0.35f, 0.3f, 0.25f))});
    const int count = std::max(1, int(std::lround(512 * detail)));
    auto S = draw_spheres(r, count, 1.0f, 0.5f, 3.0f);
    for (auto& s : S) {
        const float x = s.center.x * 90.0f, z = s.center.z * 90.0f;
        s.center = Vector3f(x, height(x, z) + 4.0f + 8.0f * (s.center.y + 1.0f), z);
    }
    add_spheres(c.builder, S, noise, 1.0f);
    add_ceiling_lights(c.builder, r, 256, 90.0f, 40.0f, 60.0f, 3.0f, 1.0f, 50.0f);
    auto sun = create_directional_light(Vector3f(0.3f, 1.0f, 0.2f), Spectrum(3.0f));
    c.builder.owned.push_back(sun);
    c.builder.lights.push_back(sun.get());
    c.width = 1920;
    c.height = 1080;
    c.spp = 256;
    c.max_depth = 32;
    c.min_depth = 3;
    c.max_leaf = 4;
    c.filter = create_box_filter(0.5f);
    auto cam = create_thin_lens_camera(1920.0f / 1080.0f, Point3f(0, 45, 150), Point3f(0, 0, 0), Vector3f(0, 1, 0),
                                       40.0f * PI_r / 180.0f, 0.0f, 1.0f);
    finish(c, cam, nullptr);
}

// A scene from an OBJ file, wired the way run() does it (src/main.cpp:212-282):
// every MTL material becomes a Phong material (a glass when Ni > 1),
// emissive materials make their triangles area lights, all
// triangles get MediumInterface{inside = vacuum, outside = fog} with the fog
// of main.cpp:242 (or vacuum outside when fog is false), and the BVH has 3
// primitives per leaf.
struct GridFog {   // a heterogeneous medium for obj_scene's outside
    const float* density = nullptr;
    int nx = 0, ny = 0, nz = 0;
    float local_to_world[12];
    float albedo[3];
    float g = 0;
    const float* albedo_grid = nullptr;   // grey albedo per voxel (same size as the density), or NULL
    const float* g_grid = nullptr;        // g per voxel, or NULL
    int max_scattering_count = 0;
};

static void obj_scene(Config& c, const std::string& path, const tr_camera& tc, bool fog, const GridFog* grid = nullptr) {
    const model_t model = load_model_from_file(path);
    std::vector<MaterialTexture> res;
    std::vector<RC<Material>> mats;
    for (const material_t& m : model.material) {
        res.push_back(create_texture_from_file(m));
        const MaterialTexture& t = res.back();
        mats.push_back(t.as_glass ? create_glass(t.map_ks, t.map_kt, t.ior)
                                  : create_phong_material(t.map_ka, t.map_kd, t.map_ks, t.map_ns));
    }
    Builder& b = c.builder;
    b.mi.inside = create_vacuum();
    b.mi.outside = fog ? create_homogeneous_medium(Spectrum(0.03f), Spectrum(0.3f), 0.8f, 25) : create_vacuum();
    if (grid) {   // create_heterogeneous_medium over an Image3D<real> (heterogeneous.cpp:101-108)
        auto im = newRC<Image3D<real>>(grid->nx, grid->ny, grid->nz);
        std::copy(grid->density, grid->density + size_t(grid->nx) * size_t(grid->ny) * size_t(grid->nz),
                  im->get_raw_data());
        real m[4][4] = {};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) m[i][j] = grid->local_to_world[4 * i + j];
        m[3][3] = 1;
        const size_t n = size_t(grid->nx) * size_t(grid->ny) * size_t(grid->nz);
        auto image_of = [&](const float* v) {
            auto img = newRC<Image3D<real>>(grid->nx, grid->ny, grid->nz);
            std::copy(v, v + n, img->get_raw_data());
            return create_image_texture3d(img);
        };
        b.mi.outside = create_heterogeneous_medium(
            Transform(m), create_image_texture3d(im),
            grid->albedo_grid ? image_of(grid->albedo_grid)
                              : create_constant_texture3d(Spectrum(grid->albedo[0], grid->albedo[1], grid->albedo[2])),
            grid->g_grid ? image_of(grid->g_grid) : create_constant_texture3d(Spectrum(grid->g)),
            grid->max_scattering_count, false);
    }
    for (const mesh_t& mesh : model.mesh) {
        auto tris = create_triangle_mesh(mesh, Transform());
        for (size_t i = 0; i < tris.size(); ++i) {
            const int id = mesh.materials[i];
            // main.cpp:263 asserts a valid material per triangle
            if (id < 0 || size_t(id) >= mats.size())
                throw std::runtime_error("OBJ scene: triangle " + std::to_string(i) + " of mesh '" + mesh.name + "' has no material");
            const Spectrum e = res[size_t(id)].has_emission ? res[size_t(id)].map_ke->evaluate(Point2f()) : Spectrum(0);
            auto p = create_geometric_primitive(tris[i], mats[size_t(id)], b.mi, e);
            if (auto l = p->as_area_light()) b.lights.push_back(l);
            b.prims.push_back(std::move(p));
        }
    }
    c.max_leaf = 3;
    c.filter = create_gaussin_filter(0.5f, 0.6f);
    auto cam = create_thin_lens_camera(tc.aspect, Point3f(tc.pos[0], tc.pos[1], tc.pos[2]),
                                       Point3f(tc.target[0], tc.target[1], tc.target[2]),
                                       Vector3f(tc.up[0], tc.up[1], tc.up[2]), tc.fov_y, tc.lens_radius,
                                       tc.focal_distance);
    finish(c, cam, nullptr);
}

static void build(Config& c, int config, uint64_t seed, double detail) {
    if (!(detail > 0.0 && detail <= 1.0)) throw std::runtime_error("detail must lie in (0, 1]");
    switch (config) {
        case 1: config1(c, seed); break;
        case 2: config2(c, seed, detail); break;
        case 3: config3(c, seed, detail); break;
        case 4: config4(c, seed, detail); break;
        case 5: config5(c, seed, detail); break;
        default: throw std::runtime_error("config must be 1..5");
    }
}

BuiltScene build_config(int config, uint64_t seed, double detail) {
    Config c;
    build(c, config, seed, detail);
    BuiltScene b;
    b.scene = c.scene;
    b.filter = c.filter;
    b.width = c.width;
    b.height = c.height;
    b.spp = c.spp;
    b.min_depth = c.min_depth;
    b.max_depth = c.max_depth;
    return b;
}

}  // namespace scenes
}  // namespace tracer

struct trs_scene {
    tracer::detail::FlatScene flat;
    tr_scene_desc desc{};
    tr_render_params defaults{};
    int max_leaf = 3;
};

static trs_scene* make_trs(tracer::scenes::Config& c, uint64_t seed);

extern "C" {

trs_scene* trs_create(int32_t config, uint64_t seed, double detail, const char** err) {
    static thread_local std::string msg;
    try {
        tracer::scenes::Config c;
        tracer::scenes::build(c, config, seed, detail);
        return make_trs(c, seed);
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return nullptr;
    }
}

trs_scene* trs_load_obj(const char* path, const tr_camera* camera, int32_t outside_fog, const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !camera) throw std::runtime_error("trs_load_obj: path and camera must not be NULL");
        tracer::scenes::Config c;
        tracer::scenes::obj_scene(c, path, *camera, outside_fog != 0);
        return make_trs(c, 0);
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return nullptr;
    }
}
trs_scene* trs_load_obj_grid(const char* path, const tr_camera* camera, const float* density, int32_t nx, int32_t ny,
                             int32_t nz, const float* local_to_world, const float* albedo, float g,
                             const float* albedo_grid, const float* g_grid, int32_t max_scattering_count,
                             const char** err) {
    static thread_local std::string msg;
    try {
        if (!path || !camera || !density || !local_to_world || !albedo)
            throw std::runtime_error("trs_load_obj_grid: pointers must not be NULL");
        if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("trs_load_obj_grid: grid dimensions must be positive");
        tracer::scenes::GridFog gf;
        gf.density = density;
        gf.nx = nx; gf.ny = ny; gf.nz = nz;
        std::copy(local_to_world, local_to_world + 12, gf.local_to_world);
        std::copy(albedo, albedo + 3, gf.albedo);
        gf.g = g;
        gf.albedo_grid = albedo_grid;
        gf.g_grid = g_grid;
        gf.max_scattering_count = max_scattering_count;
        tracer::scenes::Config c;
        tracer::scenes::obj_scene(c, path, *camera, false, &gf);
        return make_trs(c, 0);
    } catch (const std::exception& e) {
        msg = e.what();
        if (err) *err = msg.c_str();
        return nullptr;
    }
}
const tr_scene_desc* trs_desc(const trs_scene* s) { return s ? &s->desc : nullptr; }
void trs_render_defaults(const trs_scene* s, tr_render_params* out) {
    if (s && out) *out = s->defaults;
}
int32_t trs_max_leaf(const trs_scene* s) { return s ? s->max_leaf : 3; }
void trs_destroy(trs_scene* s) { delete s; }

}  // extern "C"

// Flattens a built scene and records its render defaults; throws on
// content the GPU path does not support.
static trs_scene* make_trs(tracer::scenes::Config& c, uint64_t seed) {
    auto s = std::make_unique<trs_scene>();
    s->flat = tracer::detail::flatten(*c.scene);
    s->desc = s->flat.desc();
    s->max_leaf = c.max_leaf;
    tr_render_params& p = s->defaults;
    p.width = c.width;
    p.height = c.height;
    p.spp = c.spp;
    p.min_depth = c.min_depth;
    p.max_depth = c.max_depth;
    p.direct_light_sample_num = 1;
    if (auto* g = dynamic_cast<const tracer::detail::GaussianFilter*>(c.filter.get())) {
        p.filter_type = TR_FILTER_GAUSSIAN;
        p.filter_radius = g->r;
        p.filter_alpha = g->alpha;
    } else {
        p.filter_type = TR_FILTER_BOX;
        p.filter_radius = c.filter->radius();
    }
    p.seed = seed;
    p.guide_fraction = 0.5f;
    p.guide_lr = 1e-3f;
    return s.release();
}
