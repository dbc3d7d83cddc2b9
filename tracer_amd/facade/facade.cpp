// Implementation of the tracer:: factory API (include/tracer/tracer.hpp) and
// of the GPU-backed PathTraceRenderer.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "facade_impl.h"

namespace tracer {

using namespace detail;

// ---- Transform (src/utility/transform.cpp) --------------------------------
static void identity(real m[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m[i][j] = i == j ? 1.0f : 0.0f;
}
static bool invert4(const real a[4][4], real out[4][4]) {
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? a[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
        if (m[piv][c] == 0.0) return false;
        if (piv != c)
            for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[piv][j]);
        const double inv = 1.0 / m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = m[r][c];
            for (int j = 0; j < 8; ++j) m[r][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i][j] = real(m[i][j + 4]);
    return true;
}
Transform::Transform() {
    identity(m);
    identity(inv);
}
Transform::Transform(const real mat[4][4]) {
    std::memcpy(m, mat, sizeof m);
    if (!invert4(m, inv)) throw std::runtime_error("Transform: singular matrix");
}
Transform::Transform(const real mat[4][4], const real inv_mat[4][4]) {
    std::memcpy(m, mat, sizeof m);
    std::memcpy(inv, inv_mat, sizeof inv);
}
Point3f Transform::operator()(const Point3f& p) const {
    const real x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    const real y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    const real z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    const real w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    if (w == 1.0f) return {x, y, z};
    return {x / w, y / w, z / w};
}
Vector3f Transform::operator()(const Vector3f& v) const {
    return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
            m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
}
Normal3f Transform::operator()(const Normal3f& n) const {   // transform.hpp:83-88 (inverse transpose)
    return {inv[0][0] * n.x + inv[1][0] * n.y + inv[2][0] * n.z, inv[0][1] * n.x + inv[1][1] * n.y + inv[2][1] * n.z,
            inv[0][2] * n.x + inv[1][2] * n.y + inv[2][2] * n.z};
}
Transform operator*(const Transform& a, const Transform& b) {
    real m[4][4], inv[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            real s = 0, t = 0;
            for (int k = 0; k < 4; ++k) {
                s += a.m[i][k] * b.m[k][j];
                t += b.inv[i][k] * a.inv[k][j];
            }
            m[i][j] = s;
            inv[i][j] = t;
        }
    return Transform(m, inv);
}
Transform inverse(const Transform& t) { return Transform(t.inv, t.m); }
Transform translate(const Vector3f& d) {
    real m[4][4], inv[4][4];
    identity(m);
    identity(inv);
    m[0][3] = d.x; m[1][3] = d.y; m[2][3] = d.z;
    inv[0][3] = -d.x; inv[1][3] = -d.y; inv[2][3] = -d.z;
    return Transform(m, inv);
}
Transform scale(real x, real y, real z) {
    real m[4][4], inv[4][4];
    identity(m);
    identity(inv);
    m[0][0] = x; m[1][1] = y; m[2][2] = z;
    inv[0][0] = 1 / x; inv[1][1] = 1 / y; inv[2][2] = 1 / z;
    return Transform(m, inv);
}
static Transform rot(int axis, real th) {
    const real s = std::sin(th), c = std::cos(th);
    real m[4][4];
    identity(m);
    const int a = (axis + 1) % 3, b = (axis + 2) % 3;
    m[a][a] = c; m[a][b] = -s;
    m[b][a] = s; m[b][b] = c;
    real t[4][4];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) t[i][j] = m[j][i];
    return Transform(m, t);
}
Transform rotate_x(real th) { return rot(0, th); }
Transform rotate_y(real th) { return rot(1, th); }   // transform.cpp:147-153
Transform rotate_z(real th) { return rot(2, th); }
Transform look_at(const Point3f& pos, const Point3f& target, const Vector3f& up) {   // transform.cpp:96-124
    const Vector3f dir = (target - pos).normalize();
    const Vector3f right = cross(up.normalize(), dir).normalize();
    const Vector3f nup = cross(dir, right);
    real c2w[4][4] = {{right.x, nup.x, dir.x, pos.x}, {right.y, nup.y, dir.y, pos.y}, {right.z, nup.z, dir.z, pos.z}, {0, 0, 0, 1}};
    Transform cw(c2w);
    return inverse(cw);
}

// ---- factories -------------------------------------------------------------
RC<Texture2D> create_constant_texture2d(const Spectrum& c) { return newRC<ConstantTexture2D>(c); }

RC<Material> create_disney(RC<const Texture2D> base_color, RC<const Texture2D> metallic, RC<const Texture2D> roughness,
                           RC<const Texture2D> transmission, RC<const Texture2D> transmission_roughness,
                           RC<const Texture2D> ior, RC<const Texture2D> specular_scale, RC<const Texture2D> specular_tint,
                           RC<const Texture2D> anisotropic, RC<const Texture2D> sheen, RC<const Texture2D> sheen_tint,
                           RC<const Texture2D> clearcoat, RC<const Texture2D> clearcoat_gloss,
                           Box<const NormalMapper> normal_mapper, RC<const BSSRDFSurface> bssrdf) {
    // disney.cpp:621,638 dereference both
    if (!normal_mapper) throw std::runtime_error("create_disney: normal_mapper must not be null");
    (void)bssrdf;
    auto m = newRC<DisneyMaterial>();
    m->normal_map = normal_mapper->map();
    m->base_color = base_color; m->metallic = metallic; m->roughness = roughness; m->transmission = transmission;
    m->transmission_roughness = transmission_roughness; m->ior = ior; m->specular_scale = specular_scale;
    m->specular_tint = specular_tint; m->anisotropic = anisotropic; m->sheen = sheen; m->sheen_tint = sheen_tint;
    m->clearcoat = clearcoat; m->clearcoat_gloss = clearcoat_gloss;
    const RC<const Texture2D>* all[] = {&m->base_color, &m->metallic, &m->roughness, &m->transmission,
                                        &m->transmission_roughness, &m->ior, &m->specular_scale, &m->specular_tint,
                                        &m->anisotropic, &m->sheen, &m->sheen_tint, &m->clearcoat, &m->clearcoat_gloss};
    for (auto* t : all)
        if (!*t) throw std::runtime_error("create_disney: texture must not be null");
    return m;
}

RC<Material> create_phong_material(RC<const Texture2D> ka, RC<const Texture2D> kd, RC<const Texture2D> ks,
                                   RC<const Texture2D> ns) {
    if (!kd || !ks || !ns) throw std::runtime_error("create_phong_material: texture must not be null");
    auto m = newRC<PhongMaterial>();
    m->ka = ka; m->kd = kd; m->ks = ks; m->ns = ns;
    return m;
}

RC<Material> create_glass(RC<const Texture2D> kr, RC<const Texture2D> kt, RC<const Texture2D> ior) {
    if (!kr || !kt || !ior) throw std::runtime_error("create_glass: texture must not be null");
    auto m = newRC<GlassMaterial>();
    m->kr = kr; m->kt = kt; m->ior = ior;
    return m;
}

// triangle.cpp:216-236
std::vector<RC<Shape>> create_triangle_mesh(const mesh_t& mesh, const Transform& t) {
    if (mesh.indices.size() % 3 != 0) throw std::runtime_error("create_triangle_mesh: index count not a multiple of 3");
    auto tm = newRC<TriangleMesh>();
    const size_t nv = mesh.vertices.size();
    if (!nv || mesh.indices.empty()) throw std::runtime_error("create_triangle_mesh: empty mesh");
    tm->p.resize(nv);
    tm->n.resize(nv);
    tm->uv.resize(nv);
    for (size_t i = 0; i < nv; ++i) {
        tm->p[i] = t(mesh.vertices[i].pos);
        tm->n[i] = t(mesh.vertices[i].n).normalize();
        tm->uv[i] = mesh.vertices[i].uv;
    }
    for (int v : mesh.indices)
        if (v < 0 || size_t(v) >= nv) throw std::runtime_error("create_triangle_mesh: index out of range");
    tm->indices = mesh.indices;
    std::vector<RC<Shape>> out;
    out.reserve(mesh.indices.size() / 3);
    for (size_t i = 0; i < mesh.indices.size() / 3; ++i) out.push_back(newRC<TriangleShape>(tm, int(i)));
    return out;
}

RC<Medium> create_homogeneous_medium(const Spectrum& sa, const Spectrum& ss, real g, int max_sc) {
    if (!(g > -1 && g < 1)) throw std::runtime_error("create_homogeneous_medium: g must lie in (-1, 1)");
    auto m = newRC<HomogeneousMedium>();
    m->sigma_a = sa; m->sigma_s = ss; m->g = g; m->max_scattering_count = max_sc;
    return m;
}
// Texture3D::evaluate (texture.hpp:144-160): wrap_repeat on each coordinate,
// then LinearSampler::Sample3D (sampler.hpp:47-67)
Spectrum ImageTexture3D::evaluate(const Point3f& p) const {
    const Image3D<real>& im = *image;
    auto wrap = [](double x) { return std::clamp(x - std::floor(x), 0.0, 1.0); };
    const int W = im.width(), H = im.height(), D = im.depth();
    const double u = wrap(p.x) * (W - 1), v = wrap(p.y) * (H - 1), k = wrap(p.z) * (D - 1);
    const int u0 = std::clamp(int(u), 0, W - 1), u1 = std::clamp(u0 + 1, 0, W - 1);
    const int v0 = std::clamp(int(v), 0, H - 1), v1 = std::clamp(v0 + 1, 0, H - 1);
    const int k0 = std::clamp(int(k), 0, D - 1), k1 = std::clamp(k0 + 1, 0, D - 1);
    const double du = u - u0, dv = v - v0, dk = k - k0;
    auto at = [&](int x, int y, int z) { return double(im.at(x, y, z)); };
    const double r = ((at(u0, v0, k0) * (1 - du) + at(u1, v0, k0) * du) * (1 - dv) +
                      (at(u0, v1, k0) * (1 - du) + at(u1, v1, k0) * du) * dv) * (1 - dk) +
                     ((at(u0, v0, k1) * (1 - du) + at(u1, v0, k1) * du) * (1 - dv) +
                      (at(u0, v1, k1) * (1 - du) + at(u1, v1, k1) * du) * dv) * dk;
    return Spectrum(real(r));
}
RC<Texture3D> create_constant_texture3d(const Spectrum& c) {
    auto t = newRC<ConstantTexture3D>();
    t->value = c;
    return t;
}
RC<Texture3D> create_image_texture3d(const RC<Image3D<real>>& image) {
    if (!image || image->width() < 1 || image->height() < 1 || image->depth() < 1)
        throw std::runtime_error("create_image_texture3d: empty image");
    auto t = newRC<ImageTexture3D>();
    t->image = image;
    const size_t n = size_t(image->width()) * size_t(image->height()) * size_t(image->depth());
    t->max_value = 0;
    for (size_t i = 0; i < n; ++i) t->max_value = std::max(t->max_value, image->get_raw_data()[i]);
    return t;
}
RC<Medium> create_heterogeneous_medium(const Transform& local_to_world, RC<const Texture3D> density,
                                       RC<const Texture3D> albedo, RC<const Texture3D> g, int max_sc, bool white) {
    if (!density || !albedo || !g) throw std::runtime_error("create_heterogeneous_medium: texture must not be null");
    for (const Texture3D* t : {density.get(), albedo.get(), g.get()})
        if (!t->is_constant() && !dynamic_cast<const ImageTexture3D*>(t))
            throw std::runtime_error("create_heterogeneous_medium: textures must be constant or from create_image_texture3d");
    if (g->is_constant()) {
        const real gv = g->evaluate_s(Point3f(0, 0, 0));
        if (!(gv > -1 && gv < 1)) throw std::runtime_error("create_heterogeneous_medium: g must lie in (-1, 1)");
    }
    auto m = newRC<HeterogeneousMedium>();
    m->local_to_world = local_to_world;
    m->density = std::move(density);
    m->albedo = std::move(albedo);
    m->g = std::move(g);
    m->max_scattering_count = max_sc;
    m->white_for_indirect = white;
    return m;
}
RC<Medium> create_vacuum() {   // vacuum.cpp:26-29: one shared instance
    static RC<Medium> v = newRC<VacuumMedium>();
    return v;
}

// environment_light.cpp:198-203: a constant texture or an image from
// create_texture2d_from_file; world_to_light is the rotation part of t.
RC<EnvironmentLight> create_ibl_light(RC<const Texture2D> tex, const Transform& t) {
    if (!tex) throw std::runtime_error("create_ibl_light: texture must not be null");
    if (!tex->is_constant() && !dynamic_cast<const ImageTexture2D*>(tex.get()))
        throw std::runtime_error("create_ibl_light: the GPU path takes a constant texture or an .hdr image");
    auto l = newRC<IBLConstant>();
    l->tex = tex;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) l->world_to_light[i][j] = t.m[i][j];
    return l;
}
RC<Light> create_directional_light(const Vector3f& dir, const Spectrum& E) {
    auto l = newRC<DirectionalLight>();
    l->direction = dir.normalize();
    l->irradiance = E;
    return l;
}

RC<Primitive> create_geometric_primitive(const RC<Shape>& shape, const RC<Material>& material, const MediumInterface& mi,
                                         const Spectrum& emission) {
    if (!mi.inside || !mi.outside) throw std::runtime_error("create_geometric_primitive: MediumInterface members must be set");
    if (!std::dynamic_pointer_cast<TriangleShape>(shape)) throw std::runtime_error("create_geometric_primitive: triangle shapes only");
    if (!material) throw std::runtime_error("create_geometric_primitive: material must not be null");
    auto p = newRC<GeometricPrimitive>();
    p->shape = shape;
    p->material = material;
    p->mi = mi;
    p->emission = emission;
    if (!emission.is_back()) {   // geometry.cpp:39-42
        p->light = newBox<DiffuseLight>();
        p->light->emission = emission;
    }
    return p;
}

RC<Aggregate> create_bvh_accel(int max_leaf) { return newRC<BVHAccel>(max_leaf); }

RC<Camera> create_thin_lens_camera(real aspect, const Point3f& pos, const Point3f& target, const Vector3f& up, real fov,
                                   real lens_radius, real focal_dist) {
    auto c = newRC<ThinLensCamera>();
    c->aspect = aspect; c->pos = pos; c->target = target; c->up = up; c->fov = fov;
    c->lens_radius = lens_radius; c->focal_dist = focal_dist;
    return c;
}
RC<Filter> create_gaussin_filter(real radius, real alpha) {
    auto f = newRC<GaussianFilter>();
    f->r = radius; f->alpha = alpha;
    return f;
}
RC<Filter> create_box_filter(real radius) {
    auto f = newRC<BoxFilter>();
    f->r = radius;
    return f;
}
namespace detail {
uint64_t next_version() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}
}  // namespace detail

RC<Scene> create_general_scene(const RC<Aggregate>& accel) {
    if (!std::dynamic_pointer_cast<BVHAccel>(accel)) throw std::runtime_error("create_general_scene: aggregate must come from create_bvh_accel");
    return newRC<GeneralScene>(accel);
}

// ---- flattening --------------------------------------------------------------
namespace detail {

tr_scene_desc FlatScene::desc() const {
    tr_scene_desc d{};
    d.num_vertices = int64_t(P.size() / 3);
    d.positions = P.data();
    d.normals = N.empty() ? nullptr : N.data();
    d.uvs = UV.empty() ? nullptr : UV.data();
    d.num_triangles = int64_t(mat.size());
    d.indices = idx.data();
    d.material_ids = mat.data();
    d.emission = emission.empty() ? nullptr : emission.data();
    d.medium_inside = med_in.empty() ? nullptr : med_in.data();
    d.medium_outside = med_out.empty() ? nullptr : med_out.data();
    d.num_materials = int32_t(mats.size());
    d.num_media = int32_t(media.size());
    d.materials = mats.data();
    d.media = media.empty() ? nullptr : media.data();
    d.has_env = has_env;
    std::memcpy(d.env_radiance, env, sizeof env);
    d.has_sun = has_sun;
    std::memcpy(d.sun_direction, sun_dir, sizeof sun_dir);
    std::memcpy(d.sun_irradiance, sun_E, sizeof sun_E);
    d.camera = cam;
    d.env_width = env_w;
    d.env_height = env_h;
    d.env_pixels = env_pix.empty() ? nullptr : env_pix.data();
    std::memcpy(d.env_world_to_light, env_rot, sizeof env_rot);
    tex_desc.clear();
    for (const ImageTexture2D& t : tex) {
        tr_texture td{};
        td.width = t.w;
        td.height = t.h;
        td.rgb8 = t.is8() ? t.rgb8.data() : nullptr;
        td.rgbf = t.is8() ? nullptr : t.rgb.data();
        tex_desc.push_back(td);
    }
    d.num_textures = int32_t(tex_desc.size());
    d.textures = tex_desc.empty() ? nullptr : tex_desc.data();
    return d;
}

static float tex_s(const RC<const Texture2D>& t) { return t->evaluate_s(Point2f(0, 0)); }
static void tex_rgb(const RC<const Texture2D>& t, float out[3]) {
    const Spectrum s = t->evaluate(Point2f(0, 0));
    out[0] = s.r; out[1] = s.g; out[2] = s.b;
}

// An image texture's id in the flattened scene (1 + index), registering it once.
static int32_t texture_id(const Texture2D* t, FlatScene& f, std::unordered_map<const Texture2D*, int32_t>& ids) {
    const auto* img = dynamic_cast<const ImageTexture2D*>(t);
    if (!img) throw std::runtime_error("render: unsupported texture type");
    auto it = ids.find(t);
    if (it != ids.end()) return it->second;
    f.tex.push_back(*img);
    const int32_t id = int32_t(f.tex.size());
    ids[t] = id;
    return id;
}

static tr_material convert_material(const Material* m, FlatScene& f, std::unordered_map<const Texture2D*, int32_t>& tex_ids) {
    tr_material r{};
    if (auto d = dynamic_cast<const DisneyMaterial*>(m)) {
        // every parameter a constant or an image read per hit (tr_material::param_texture
        // order), and the normal map of the NormalMapper
        r.type = TR_MAT_DISNEY;
        if (d->base_color->is_constant()) tex_rgb(d->base_color, r.base_color);
        else r.base_color_texture = texture_id(d->base_color.get(), f, tex_ids);
        const RC<const Texture2D>* params[TR_PARAM_COUNT] = {
            &d->metallic, &d->roughness, &d->transmission, &d->transmission_roughness, &d->ior, &d->specular_scale,
            &d->specular_tint, &d->anisotropic, &d->sheen, &d->sheen_tint, &d->clearcoat, &d->clearcoat_gloss};
        float* consts[TR_PARAM_COUNT] = {&r.metallic, &r.roughness, &r.transmission, &r.transmission_roughness,
                                         &r.ior, nullptr, &r.specular_tint, &r.anisotropic, &r.sheen,
                                         &r.sheen_tint, &r.clearcoat, &r.clearcoat_gloss};
        for (int k = 0; k < TR_PARAM_COUNT; ++k) {
            const RC<const Texture2D>& t = *params[k];
            if (!t->is_constant()) {
                r.param_texture[k] = texture_id(t.get(), f, tex_ids);
            } else if (k == TR_PARAM_SPECULAR_SCALE) {
                tex_rgb(t, r.specular_scale);
            } else {
                *consts[k] = tex_s(t);
            }
        }
        if (d->normal_map) r.normal_texture = texture_id(d->normal_map.get(), f, tex_ids);
        return r;
    }
    if (auto p = dynamic_cast<const PhongMaterial*>(m)) {
        if (!p->ks->is_constant() || !p->ns->is_constant())
            throw std::runtime_error("Phong: only kd may be an image texture on the GPU path");
        if (!p->kd->is_constant()) {   // kd per hit from the image; the library normalises per hit
            const Spectrum ks = p->ks->evaluate(Point2f(0, 0));
            r.type = TR_MAT_PHONG;
            r.base_color_texture = texture_id(p->kd.get(), f, tex_ids);
            r.specular_scale[0] = ks.r; r.specular_scale[1] = ks.g; r.specular_scale[2] = ks.b;
            r.shininess = tex_s(p->ns);
            return r;
        }
        const Spectrum kd = p->kd->evaluate(Point2f(0, 0)), ks = p->ks->evaluate(Point2f(0, 0));
        if (!ks.is_back()) {
            // diffuse + Phong specular lobe; the library normalises and weighs them
            r.type = TR_MAT_PHONG;
            r.base_color[0] = kd.r; r.base_color[1] = kd.g; r.base_color[2] = kd.b;
            r.specular_scale[0] = ks.r; r.specular_scale[1] = ks.g; r.specular_scale[2] = ks.b;
            r.shininess = tex_s(p->ns);
            return r;
        }
        // ks = 0: a single diffuse lobe.  phong.cpp:36-41: energy conservation
        // divides by max(1, max_i(kd_i + ks_i))
        const float dem = std::max({1.0f, kd.r + ks.r, kd.g + ks.g, kd.b + ks.b});
        r.type = TR_MAT_LAMBERT;
        r.base_color[0] = kd.r / dem; r.base_color[1] = kd.g / dem; r.base_color[2] = kd.b / dem;
        return r;
    }
    if (auto g = dynamic_cast<const GlassMaterial*>(m)) {
        if (!g->kr->is_constant() || !g->kt->is_constant() || !g->ior->is_constant())
            throw std::runtime_error("Glass: only constant textures are supported");
        r.type = TR_MAT_GLASS;
        tex_rgb(g->kr, r.base_color);
        tex_rgb(g->kt, r.specular_scale);
        r.ior = tex_s(g->ior);
        return r;
    }
    throw std::runtime_error("unsupported material");
}

FlatScene flatten(const Scene& scene) {
    const auto* gs = dynamic_cast<const GeneralScene*>(&scene);
    if (!gs) throw std::runtime_error("render: scene must come from create_general_scene");
    const auto* bvh = dynamic_cast<const BVHAccel*>(gs->accel.get());
    if (!bvh || bvh->prims.empty()) throw std::runtime_error("render: aggregate has no primitives (call build)");
    const auto* cam = dynamic_cast<const ThinLensCamera*>(gs->camera.get());
    if (!cam) throw std::runtime_error("render: scene has no camera");
    FlatScene f;
    f.max_leaf = bvh->max_leaf;
    std::unordered_map<const TriangleMesh*, int32_t> mesh_base;
    std::unordered_map<const Material*, int32_t> mat_id;
    std::unordered_map<const Texture2D*, int32_t> tex_ids;
    std::unordered_map<const Medium*, int32_t> med_id;
    auto medium_index = [&](const RC<const Medium>& m) -> int32_t {
        if (dynamic_cast<const VacuumMedium*>(m.get())) return -1;
        auto it = med_id.find(m.get());
        if (it != med_id.end()) return it->second;
        tr_medium tm{};
        if (const auto* hm = dynamic_cast<const HeterogeneousMedium*>(m.get())) {
            // albedo and g: constants, or grids looked up per collision (a grey
            // image gives the same albedo in every channel)
            if (const auto* ia = dynamic_cast<const ImageTexture3D*>(hm->albedo.get())) {
                const Image3D<real>& im = *ia->image;
                tm.anx = im.width(); tm.any = im.height(); tm.anz = im.depth();
                const size_t n = size_t(tm.anx) * size_t(tm.any) * size_t(tm.anz);
                f.grids.emplace_back(3 * n);
                for (size_t k = 0; k < n; ++k)
                    for (int c = 0; c < 3; ++c) f.grids.back()[3 * k + size_t(c)] = im.get_raw_data()[k];
                tm.albedo_grid = f.grids.back().data();
                tm.sigma_s[0] = tm.sigma_s[1] = tm.sigma_s[2] = 1.0f;
            } else {
                const Spectrum a = hm->albedo->evaluate(Point3f(0, 0, 0));
                tm.sigma_s[0] = a.r; tm.sigma_s[1] = a.g; tm.sigma_s[2] = a.b;
            }
            if (const auto* ig = dynamic_cast<const ImageTexture3D*>(hm->g.get())) {
                const Image3D<real>& im = *ig->image;
                tm.gnx = im.width(); tm.gny = im.height(); tm.gnz = im.depth();
                f.grids.emplace_back(im.get_raw_data(), im.get_raw_data() + size_t(tm.gnx) * size_t(tm.gny) * size_t(tm.gnz));
                tm.g_grid = f.grids.back().data();
                tm.g = 0.0f;
            } else {
                tm.g = hm->g->evaluate_s(Point3f(0, 0, 0));
            }
            tm.max_scattering_count = hm->max_scattering_count;
            tm.white_for_indirect = hm->white_for_indirect ? 1 : 0;
            // the flat scene keeps its own copy of the voxels (a constant density
            // is a grid of one voxel)
            if (const auto* it = dynamic_cast<const ImageTexture3D*>(hm->density.get())) {
                const Image3D<real>& im = *it->image;
                tm.nx = im.width(); tm.ny = im.height(); tm.nz = im.depth();
                f.grids.emplace_back(im.get_raw_data(), im.get_raw_data() + size_t(tm.nx) * size_t(tm.ny) * size_t(tm.nz));
            } else {
                tm.nx = tm.ny = tm.nz = 1;
                f.grids.emplace_back(1, hm->density->evaluate_s(Point3f(0, 0, 0)));
            }
            tm.density = f.grids.back().data();
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 4; ++j) tm.local_to_world[4 * i + j] = hm->local_to_world.m[i][j];
            const int32_t id = int32_t(f.media.size());
            f.media.push_back(tm);
            med_id[m.get()] = id;
            return id;
        }
        const auto* h = dynamic_cast<const HomogeneousMedium*>(m.get());
        if (!h) throw std::runtime_error("render: only homogeneous and heterogeneous media and vacuum are supported");
        tm.sigma_a[0] = h->sigma_a.r; tm.sigma_a[1] = h->sigma_a.g; tm.sigma_a[2] = h->sigma_a.b;
        tm.sigma_s[0] = h->sigma_s.r; tm.sigma_s[1] = h->sigma_s.g; tm.sigma_s[2] = h->sigma_s.b;
        tm.g = h->g;
        tm.max_scattering_count = h->max_scattering_count;
        const int32_t id = int32_t(f.media.size());
        f.media.push_back(tm);
        med_id[m.get()] = id;
        return id;
    };
    const size_t np = bvh->prims.size();
    f.idx.reserve(3 * np);
    f.mat.reserve(np);
    f.med_in.reserve(np);
    f.med_out.reserve(np);
    f.emission.reserve(3 * np);
    bool any_medium = false, any_emission = false;
    for (const auto& prim : bvh->prims) {
        const auto* gp = dynamic_cast<const GeometricPrimitive*>(prim.get());
        if (!gp) throw std::runtime_error("render: unsupported primitive");
        const auto* tri = static_cast<const TriangleShape*>(gp->shape.get());
        const TriangleMesh* mesh = tri->mesh.get();
        auto it = mesh_base.find(mesh);
        int32_t base;
        if (it == mesh_base.end()) {
            base = int32_t(f.P.size() / 3);
            for (size_t v = 0; v < mesh->p.size(); ++v) {
                f.P.insert(f.P.end(), {mesh->p[v].x, mesh->p[v].y, mesh->p[v].z});
                f.N.insert(f.N.end(), {mesh->n[v].x, mesh->n[v].y, mesh->n[v].z});
                f.UV.insert(f.UV.end(), {mesh->uv[v].x, mesh->uv[v].y});
            }
            mesh_base[mesh] = base;
        } else {
            base = it->second;
        }
        for (int k = 0; k < 3; ++k) f.idx.push_back(base + mesh->indices[3 * size_t(tri->index) + k]);
        auto mit = mat_id.find(gp->material.get());
        int32_t mid;
        if (mit == mat_id.end()) {
            mid = int32_t(f.mats.size());
            f.mats.push_back(convert_material(gp->material.get(), f, tex_ids));
            mat_id[gp->material.get()] = mid;
        } else {
            mid = mit->second;
        }
        f.mat.push_back(mid);
        const int32_t mi = medium_index(gp->mi.inside), mo = medium_index(gp->mi.outside);
        any_medium |= (mi >= 0 || mo >= 0);
        f.med_in.push_back(mi);
        f.med_out.push_back(mo);
        f.emission.insert(f.emission.end(), {gp->emission.r, gp->emission.g, gp->emission.b});
        any_emission |= !gp->emission.is_back();
    }
    if (!any_medium) { f.med_in.clear(); f.med_out.clear(); }
    if (!any_emission) f.emission.clear();
    if (gs->environment_light) {
        const auto* ibl = dynamic_cast<const IBLConstant*>(gs->environment_light.get());
        if (!ibl) throw std::runtime_error("render: unsupported environment light");
        f.has_env = 1;
        if (const auto* img = dynamic_cast<const ImageTexture2D*>(ibl->tex.get())) {
            f.env_w = img->w;
            f.env_h = img->h;
            f.env_pix = img->rgb;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) f.env_rot[3 * i + j] = ibl->world_to_light[i][j];
        } else {
            tex_rgb(ibl->tex, f.env);
        }
    }
    for (const Light* l : gs->lights) {
        if (const auto* dl = dynamic_cast<const DirectionalLight*>(l)) {
            if (f.has_sun) throw std::runtime_error("render: at most one directional light is supported");
            f.has_sun = 1;
            f.sun_dir[0] = dl->direction.x; f.sun_dir[1] = dl->direction.y; f.sun_dir[2] = dl->direction.z;
            f.sun_E[0] = dl->irradiance.r; f.sun_E[1] = dl->irradiance.g; f.sun_E[2] = dl->irradiance.b;
        }
    }
    f.cam = camera_desc(scene);
    return f;
}

tr_camera camera_desc(const Scene& scene) {
    const auto* cam = dynamic_cast<const ThinLensCamera*>(scene.get_camera().get());
    if (!cam) throw std::runtime_error("render: scene has no camera");
    tr_camera c{};
    c.pos[0] = cam->pos.x; c.pos[1] = cam->pos.y; c.pos[2] = cam->pos.z;
    c.target[0] = cam->target.x; c.target[1] = cam->target.y; c.target[2] = cam->target.z;
    c.up[0] = cam->up.x; c.up[1] = cam->up.y; c.up[2] = cam->up.z;
    c.fov_y = cam->fov;
    c.lens_radius = cam->lens_radius;
    c.focal_distance = cam->focal_dist;
    c.aspect = cam->aspect;
    return c;
}

}  // namespace detail

// ---- the GPU path tracer behind create_pt_renderer ---------------------------
namespace {

// The device context and the uploaded scene of a GPU renderer.  The geometry
// is uploaded again for another scene object or a rebuilt aggregate, or when
// the light list or environment changed; the camera is sent with every
// render, so set_camera between renders takes effect.
class GpuRendererBase : public Renderer {
public:
    explicit GpuRendererBase(const GpuRenderOptions& g) : gpu(g) {}
    ~GpuRendererBase() override {
        if (ctx) tr_destroy(ctx);
    }

protected:
    void check(tr_status s, const char* what) {
        if (s != TR_OK) throw std::runtime_error(std::string(what) + ": " + (ctx ? tr_last_error(ctx) : "no context"));
    }
    // Call with mtx held.
    void prepare(const Scene& scene) {
        if (!ctx) check(tr_create(&ctx, gpu.device), "tr_create");
        const Key key = scene_key(scene);
        if (!(key == last_key)) {
            FlatScene f = flatten(scene);
            const tr_scene_desc d = f.desc();
            check(tr_scene_upload(ctx, &d), "tr_scene_upload");
            tr_bvh_opts bo{};
            bo.max_leaf_primitives = std::min(std::max(f.max_leaf, 1), 15);
            bo.num_threads = 0;
            check(tr_bvh_build(ctx, &bo), "tr_bvh_build");
            last_key = key;
        }
        const tr_camera cam = camera_desc(scene);
        check(tr_set_camera(ctx, &cam), "tr_set_camera");
    }
    static void film_filter(const Film& film, int32_t& type, float& radius, float& alpha) {
        const Filter* flt = film.get_filter().get();
        if (const auto* gf = dynamic_cast<const GaussianFilter*>(flt)) {
            type = TR_FILTER_GAUSSIAN;
            radius = gf->r;
            alpha = gf->alpha;
        } else if (const auto* bf = dynamic_cast<const BoxFilter*>(flt)) {
            type = TR_FILTER_BOX;
            radius = bf->r;
            alpha = 0.0f;
        } else {
            throw std::runtime_error("render: unsupported filter");
        }
    }

    GpuRenderOptions gpu;
    tr_ctx* ctx = nullptr;
    std::mutex mtx;

private:
    struct Key {
        uint64_t scene = 0, accel = 0;
        size_t lights = 0;
        const void* env = nullptr;
        bool operator==(const Key& o) const { return scene == o.scene && accel == o.accel && lights == o.lights && env == o.env; }
    };
    static Key scene_key(const Scene& s) {
        Key k;
        if (const auto* gs = dynamic_cast<const GeneralScene*>(&s)) {
            k.scene = gs->uid;
            if (const auto* b = dynamic_cast<const BVHAccel*>(gs->accel.get())) k.accel = b->version;
        }
        k.lights = s.lights.size();
        k.env = s.environment_light.get();
        return k;
    }
    Key last_key;
};

class GpuPathTraceRenderer : public GpuRendererBase {
public:
    GpuPathTraceRenderer(const PTRendererParams& p, const GpuRenderOptions& g) : GpuRendererBase(g), params(p) {
        if (p.spp <= 0 || p.max_depth < 0 || p.min_depth < 0) throw std::runtime_error("create_pt_renderer: invalid params");
    }
    RenderTarget render(const Scene& scene, Film film) override {
        std::lock_guard<std::mutex> lk(mtx);
        prepare(scene);
        tr_render_params rp{};
        rp.width = film.width();
        rp.height = film.height();
        rp.spp = params.spp;
        rp.min_depth = params.min_depth;
        rp.max_depth = params.max_depth;
        rp.direct_light_sample_num = params.direct_light_sample_num;
        film_filter(film, rp.filter_type, rp.filter_radius, rp.filter_alpha);
        rp.shadow_transmittance = gpu.shadow_transmittance;
        rp.seed = gpu.seed;
        rp.samples_per_pass = gpu.samples_per_pass;
        rp.guide = gpu.guide;
        rp.guide_train_passes = gpu.guide_train_passes;
        rp.guide_fraction = gpu.guide_fraction;
        rp.guide_lr = gpu.guide_lr;
        rp.guide_train_fraction = gpu.guide_train_fraction;
        rp.guide_max_depth = gpu.guide_max_depth;
        rp.tail_paths = gpu.tail_paths;
        rp.guide_light_grid = gpu.guide_light_grid;
        RenderTarget rt;
        rt.color = Image2D<Spectrum>(rp.width, rp.height);
        static_assert(sizeof(Spectrum) == 12, "Spectrum is three floats");
        tr_render_stats st{};
        check(tr_render(ctx, &rp, nullptr, 0, reinterpret_cast<float*>(rt.color.get_raw_data()), &st), "tr_render");
        if (gpu.stats_out) *gpu.stats_out = st;
        return rt;
    }

private:
    PTRendererParams params;
};

// BDPTRenderer (src/renderer/bdpt_renderer.cpp:1066-1185) on the GPU.
class GpuBDPTRenderer : public GpuRendererBase {
public:
    GpuBDPTRenderer(const BDPTRendererParams& p, const GpuRenderOptions& g) : GpuRendererBase(g), params(p) {
        if (p.spp <= 0 || p.max_camera_vertex_count < 1 || p.max_light_vertex_count < 0)
            throw std::runtime_error("create_bdpt_renderer: invalid params");
    }
    RenderTarget render(const Scene& scene, Film film) override {
        std::lock_guard<std::mutex> lk(mtx);
        prepare(scene);
        tr_bdpt_params bp{};
        bp.width = film.width();
        bp.height = film.height();
        bp.spp = params.spp;
        bp.max_camera_vertex_count = params.max_camera_vertex_count;
        bp.max_light_vertex_count = params.max_light_vertex_count;
        film_filter(film, bp.filter_type, bp.filter_radius, bp.filter_alpha);
        bp.seed = gpu.seed;
        bp.samples_per_pass = gpu.samples_per_pass;
        RenderTarget rt;
        rt.color = Image2D<Spectrum>(bp.width, bp.height);
        tr_render_stats st{};
        check(tr_render_bdpt(ctx, &bp, nullptr, 0, reinterpret_cast<float*>(rt.color.get_raw_data()), &st),
              "tr_render_bdpt");
        if (gpu.stats_out) *gpu.stats_out = st;
        return rt;
    }

private:
    BDPTRendererParams params;
};

}  // namespace

RC<Renderer> create_pt_renderer(const PTRendererParams& params) { return newRC<GpuPathTraceRenderer>(params, GpuRenderOptions{}); }
RC<Renderer> create_pt_renderer(const PTRendererParams& params, const GpuRenderOptions& gpu) {
    return newRC<GpuPathTraceRenderer>(params, gpu);
}
RC<Renderer> create_bdpt_renderer(const BDPTRendererParams& params) {
    return newRC<GpuBDPTRenderer>(params, GpuRenderOptions{});
}
RC<Renderer> create_bdpt_renderer(const BDPTRendererParams& params, const GpuRenderOptions& gpu) {
    return newRC<GpuBDPTRenderer>(params, gpu);
}

}  // namespace tracer
