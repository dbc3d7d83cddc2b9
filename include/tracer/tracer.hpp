// tracer.hpp — the reference's C++ factory API (namespace tracer), backed by
// the MI355X wavefront path tracer in libtracer_hip.so.
//
// Names, signatures and argument meaning follow the reference headers so a
// scene assembled the way src/main.cpp assembles one renders unchanged:
//   create_pt_renderer          src/factory/renderer.hpp:12-21
//   create_bdpt_renderer        src/factory/renderer.hpp:39-48
//   create_general_scene        src/factory/scene.hpp:14
//   create_bvh_accel            src/factory/accelerator.hpp:12
//   create_triangle_mesh        src/factory/shape.hpp:17
//   create_geometric_primitive  src/factory/primivite.hpp:12-14
//   create_disney               src/factory/material.hpp:37-52
//   create_phong_material       src/factory/material.hpp:19-24
//   create_glass                src/factory/material.hpp:13-17
//   load_model_from_file        src/utility/mesh_load.hpp:93, mesh_load.cpp:25-113
//   create_texture_from_file    src/utility/image_file.hpp:23-45, image_file.cpp:139-166 (constant colours)
//   create_constant_texture2d   src/factory/texture.hpp:12
//   create_ibl_light            src/factory/light.hpp:12-15 (constant texture or .hdr image)
//   create_texture2d_from_file  src/utility/image_file.hpp:21 (Radiance .hdr)
//   write_image_to_hdr / _png   src/utility/image_file.hpp:12-15
//   create_gamma_corrector, create_aces_tone_mapper  src/post_processor/*.cpp
//   create_homogeneous_medium   src/factory/medium.hpp:21-26
//   create_vacuum               src/factory/medium.hpp:34
//   create_thin_lens_camera     src/factory/camera.hpp:9-11
//   create_gaussin_filter       src/factory/filter.hpp:12
//   Renderer::render            src/core/renderer.hpp:16
//   Film / RenderTarget         src/core/render.hpp:16-19, 74-76
// Additions without a reference counterpart: create_box_filter,
// create_directional_light, GpuRenderOptions and the create_pt_renderer
// overload that takes them.
//
// Errors: construction errors throw std::runtime_error, as in the reference
// (caught by the caller, src/main.cpp:460-468).  Renderer::render throws
// std::runtime_error with the library's message if the GPU path fails.
#pragma once
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct tr_render_stats;

namespace tracer {

using real = float;
template <class T> using RC = std::shared_ptr<T>;
template <class T> using Box = std::unique_ptr<T>;
template <class T> using Span = std::vector<T>;
template <class T, class... A> RC<T> newRC(A&&... a) { return std::make_shared<T>(std::forward<A>(a)...); }
template <class T, class... A> Box<T> newBox(A&&... a) { return std::make_unique<T>(std::forward<A>(a)...); }

constexpr real PI_r = real(3.14159265358979323846);

struct Vector3f {
    real x = 0, y = 0, z = 0;
    Vector3f() = default;
    Vector3f(real a, real b, real c) : x(a), y(b), z(c) {}
    Vector3f operator+(const Vector3f& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3f operator-(const Vector3f& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector3f operator*(real s) const { return {x * s, y * s, z * s}; }
    Vector3f operator-() const { return {-x, -y, -z}; }
    real length() const { return std::sqrt(x * x + y * y + z * z); }
    Vector3f normalize() const { real l = length(); return l > 0 ? Vector3f(x / l, y / l, z / l) : *this; }
};
struct Point3f {
    real x = 0, y = 0, z = 0;
    Point3f() = default;
    Point3f(real a, real b, real c) : x(a), y(b), z(c) {}
    Point3f operator+(const Vector3f& v) const { return {x + v.x, y + v.y, z + v.z}; }
    Vector3f operator-(const Point3f& o) const { return {x - o.x, y - o.y, z - o.z}; }
};
struct Normal3f {
    real x = 0, y = 0, z = 0;
    Normal3f() = default;
    Normal3f(real a, real b, real c) : x(a), y(b), z(c) {}
    Normal3f normalize() const { real l = std::sqrt(x * x + y * y + z * z); return l > 0 ? Normal3f(x / l, y / l, z / l) : *this; }
};
struct Point2f {
    real x = 0, y = 0;
    Point2f() = default;
    Point2f(real a, real b) : x(a), y(b) {}
};
struct Point2i {
    int x = 0, y = 0;
    Point2i() = default;
    Point2i(int a, int b) : x(a), y(b) {}
};
inline Vector3f cross(const Vector3f& a, const Vector3f& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline real dot(const Vector3f& a, const Vector3f& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Color3<float> (src/core/spectrum.hpp:12-162)
struct Spectrum {
    real r = 0, g = 0, b = 0;
    Spectrum() = default;
    explicit Spectrum(real v) : r(v), g(v), b(v) {}
    Spectrum(real a, real c, real d) : r(a), g(c), b(d) {}
    real lum() const { return real(0.2126) * r + real(0.7152) * g + real(0.0722) * b; }
    bool is_back() const { return r == 0 && g == 0 && b == 0; }   // "is black" (spectrum.hpp)
    Spectrum operator*(real s) const { return {r * s, g * s, b * s}; }
    Spectrum operator+(const Spectrum& o) const { return {r + o.r, g + o.g, b + o.b}; }
};

// Transform (src/utility/transform.hpp:50-99): row-major 4x4 matrix and inverse.
class Transform {
public:
    real m[4][4];
    real inv[4][4];
    Transform();
    Transform(const real mat[4][4]);
    Transform(const real mat[4][4], const real inv_mat[4][4]);
    Point3f operator()(const Point3f& p) const;
    Vector3f operator()(const Vector3f& v) const;
    Normal3f operator()(const Normal3f& n) const;
    friend Transform operator*(const Transform& a, const Transform& b);
    friend Transform inverse(const Transform& t);
};
Transform translate(const Vector3f& delta);
Transform scale(real x, real y, real z);
Transform rotate_x(real theta);   // radians
Transform rotate_y(real theta);
Transform rotate_z(real theta);
Transform look_at(const Point3f& pos, const Point3f& target, const Vector3f& up);

// mesh_load.hpp:15-25, 81-89
struct vertex_t {
    Point3f pos;
    Normal3f n;
    Point2f uv;
};
struct mesh_t {
    std::string name;
    std::vector<vertex_t> vertices;
    std::vector<int> indices;
    std::vector<int> materials;
};
// mesh_load.hpp:28-79: one MTL material as the OBJ reader returns it.
struct material_t {
    std::string name;
    real ambient[3] = {0, 0, 0};
    real diffuse[3] = {0, 0, 0};
    real specular[3] = {0, 0, 0};
    real transmittance[3] = {0, 0, 0};
    real emission[3] = {0, 0, 0};
    real shininess = 1;
    real ior = 1;        // index of refraction
    real dissolve = 1;   // 1 == opaque; 0 == fully transparent
    int illum = 0;
    std::string map_ka, map_kd, map_ks, map_ns, map_bump, disp, map_d, refl;
    real roughness = 0, metallic = 0, sheen = 0, clearcoat_thickness = 0, clearcoat_roughness = 0;
    std::string map_pr, map_pm, map_ps, map_ke, norm;
};
// mesh_load.hpp:81-86
struct model_t {
    std::string name;
    std::vector<mesh_t> mesh;
    std::vector<material_t> material;
};
// Reads a Wavefront OBJ file and the MTL libraries it names (mesh_load.cpp:25-113):
// one mesh per object/group with vertices deduplicated on (position, normal, uv),
// faces triangulated, one material index per triangle (-1 = none).  Throws
// std::runtime_error when the file cannot be read or parsed.
model_t load_model_from_file(const std::string& path);

template <class T>
class Image2D {
public:
    Image2D() = default;
    Image2D(int width, int height) : w(width), h(height), data(size_t(width) * height) {}
    T& at(int x, int y) { return data[size_t(y) * w + x]; }
    const T& at(int x, int y) const { return data[size_t(y) * w + x]; }
    int width() const { return w; }
    int height() const { return h; }
    T* get_raw_data() { return data.data(); }
    const T* get_raw_data() const { return data.data(); }
private:
    int w = 0, h = 0;
    std::vector<T> data;
};

class Texture2D {
public:
    virtual ~Texture2D() = default;
    virtual Spectrum evaluate(const Point2f& uv) const = 0;
    virtual real evaluate_s(const Point2f& uv) const { return evaluate(uv).r; }
    virtual bool is_constant() const { return false; }
};
RC<Texture2D> create_constant_texture2d(const Spectrum& constant);
// image_file.hpp:21: a Radiance .hdr file (linear RGB, the texture for an
// image environment or any colour texture) or a PNG file (8-bit, rows flipped
// as the reference loads them).  Other formats throw std::runtime_error.
RC<const Texture2D> create_texture2d_from_file(const std::string& filename);

// Volumes (src/utility/image.hpp, src/core/texture.hpp:108-170,
// src/factory/texture.hpp:18-20): x fastest, then y, then z.
template <class T>
class Image3D {
public:
    Image3D() = default;
    Image3D(int width, int height, int depth)
        : w(width), h(height), d(depth), data(size_t(width) * size_t(height) * size_t(depth)) {}
    T& at(int x, int y, int z) { return data[(size_t(z) * size_t(h) + size_t(y)) * size_t(w) + size_t(x)]; }
    const T& at(int x, int y, int z) const { return data[(size_t(z) * size_t(h) + size_t(y)) * size_t(w) + size_t(x)]; }
    int width() const { return w; }
    int height() const { return h; }
    int depth() const { return d; }
    T* get_raw_data() { return data.data(); }
    const T* get_raw_data() const { return data.data(); }
private:
    int w = 0, h = 0, d = 0;
    std::vector<T> data;
};

class Texture3D {
public:
    virtual ~Texture3D() = default;
    virtual Spectrum evaluate(const Point3f& uvw) const = 0;   // wrap_repeat, trilinear
    virtual real evaluate_s(const Point3f& uvw) const { return evaluate(uvw).r; }
    virtual real max_real() const = 0;
    virtual bool is_constant() const { return false; }
};
RC<Texture3D> create_constant_texture3d(const Spectrum& constant);
RC<Texture3D> create_image_texture3d(const RC<Image3D<real>>& image);

class NormalMapper {
public:
    NormalMapper() = default;
    explicit NormalMapper(RC<const Texture2D> normal_map) noexcept : normal_map(std::move(normal_map)) {}
    const RC<const Texture2D>& map() const { return normal_map; }
private:
    RC<const Texture2D> normal_map;
};
class BSSRDFSurface {
public:
    virtual ~BSSRDFSurface() = default;
};

class Material {
public:
    virtual ~Material() = default;
};
// Disney material (src/factory/material.hpp:37-52, src/material/disney.cpp):
// every parameter a constant or an image texture read at each hit's uv
// (scalars from the red channel), and the normal map of normal_mapper, if it
// has one, turning the shading frame per hit.
RC<Material> create_disney(RC<const Texture2D> base_color, RC<const Texture2D> metallic, RC<const Texture2D> roughness,
                           RC<const Texture2D> transmission, RC<const Texture2D> transmission_roughness,
                           RC<const Texture2D> ior, RC<const Texture2D> specular_scale, RC<const Texture2D> specular_tint,
                           RC<const Texture2D> anisotropic, RC<const Texture2D> sheen, RC<const Texture2D> sheen_tint,
                           RC<const Texture2D> clearcoat, RC<const Texture2D> clearcoat_gloss,
                           Box<const NormalMapper> normal_mapper, RC<const BSSRDFSurface> bssrdf);
RC<Material> create_phong_material(RC<const Texture2D> map_ka, RC<const Texture2D> map_kd, RC<const Texture2D> map_ks,
                                   RC<const Texture2D> map_ns);
// Specular dielectric (src/material/glass.cpp): reflection colour map_kr,
// refraction colour sqrt(map_kt), index of refraction map_ior.
RC<Material> create_glass(RC<const Texture2D> map_kr, RC<const Texture2D> map_kt, RC<const Texture2D> map_ior);

// image_file.hpp:23-43: the textures of one MTL material.
struct MaterialTexture {
    RC<Texture2D> map_ka, map_kd, map_ks, map_kt, ior, map_ns, map_bump, disp, map_d, refl;
    RC<Texture2D> map_pr, map_pm, map_ps, map_ke, norm;
    bool has_emission = false;
    bool as_glass = false;
};
// image_file.cpp:139-166: Ka, Kd, Ks, Ns and Ke become constant textures,
// map_Kd an image texture (8-bit maps decoded and scaled by Kd as the
// reference does), has_emission is set when Ke > 0, and ior > 1 makes a glass
// (Ks, or white when Ks is black, as reflectance and transmittance).  Image
// maps other than map_Kd throw std::runtime_error: the GPU path samples image
// textures for the diffuse colour only.
MaterialTexture create_texture_from_file(const material_t& material);

class Shape {
public:
    virtual ~Shape() = default;
};
std::vector<RC<Shape>> create_triangle_mesh(const mesh_t& mesh, const Transform& local_to_world);

class Medium {
public:
    virtual ~Medium() = default;
};
RC<Medium> create_homogeneous_medium(const Spectrum& sigma_a, const Spectrum& sigma_s, real g, int max_scattering_count);
RC<Medium> create_vacuum();
// heterogeneous.cpp:101-108.  Density, albedo and g are each an image or a
// constant 3-D texture; an image albedo or g is looked up at every collision
// (a grey image gives the same albedo in every channel).
RC<Medium> create_heterogeneous_medium(const Transform& local_to_world, RC<const Texture3D> density,
                                       RC<const Texture3D> albedo, RC<const Texture3D> g, int max_scattering_count,
                                       bool white_for_indirect);
struct MediumInterface {
    RC<const Medium> inside;
    RC<const Medium> outside;
};

class Light {
public:
    virtual ~Light() = default;
};
class AreaLight : public Light {};
class EnvironmentLight : public Light {};
RC<EnvironmentLight> create_ibl_light(RC<const Texture2D> tex, const Transform& world_to_light);
// Directional light; `direction` points from the scene towards the light.
RC<Light> create_directional_light(const Vector3f& direction, const Spectrum& irradiance);

class Primitive {
public:
    virtual ~Primitive() = default;
    virtual const AreaLight* as_area_light() const noexcept = 0;
};
RC<Primitive> create_geometric_primitive(const RC<Shape>& shape, const RC<Material>& material, const MediumInterface& mi,
                                         const Spectrum& emission);

class Aggregate {
public:
    virtual ~Aggregate() = default;
    virtual void build(std::vector<RC<Primitive>> prims) = 0;
};
RC<Aggregate> create_bvh_accel(int max_leaf_primitives);

class Camera {
public:
    virtual ~Camera() = default;
};
RC<Camera> create_thin_lens_camera(real film_aspect, const Point3f& cam_pos, const Point3f& target, const Vector3f& up,
                                   real fov, real lens_radius, real focal_dist);

class Filter {
public:
    virtual ~Filter() = default;
    virtual real radius() const noexcept = 0;
};
RC<Filter> create_gaussin_filter(real radius, real alpha);
RC<Filter> create_box_filter(real radius);

class Scene {
public:
    virtual ~Scene() = default;
    virtual void set_camera(RC<const Camera>) = 0;
    virtual RC<const Camera> get_camera() const noexcept = 0;
    virtual void prepare_to_render() = 0;
    Span<const Light*> lights;
    RC<EnvironmentLight> environment_light;
};
RC<Scene> create_general_scene(const RC<Aggregate>& accel);

struct RenderTarget {
    Image2D<Spectrum> color;
};

// Point3b / Color3b (src/utility/geometry.hpp:635): 8-bit RGB.
struct Color3b {
    uint8_t x = 0, y = 0, z = 0;
};
// Radiance RGBE file (run-length scanlines) and 8-bit RGB PNG; throw
// std::runtime_error when the file cannot be written.
void write_image_to_hdr(const Image2D<Spectrum>& image, const std::string& filename);
void write_image_to_png(const Image2D<Color3b>& image, const std::string& filename);

// src/core/post_processor.hpp:12-17
class PostProcessor {
public:
    virtual ~PostProcessor() = default;
    virtual void process(RenderTarget& render_target) = 0;
};
RC<PostProcessor> create_gamma_corrector(real gamma);        // every channel to the power gamma
RC<PostProcessor> create_aces_tone_mapper(real exposure);    // fitted ACES curve of exposure * channel

class Film {
public:
    Film(const Point2i& res, const RC<Filter>& filter) : resolution(res), filter(filter) {}
    int width() const { return resolution.x; }
    int height() const { return resolution.y; }
    const RC<Filter>& get_filter() const { return filter; }
private:
    Point2i resolution;
    RC<Filter> filter;
};

class Renderer {
public:
    virtual ~Renderer() = default;
    virtual RenderTarget render(const Scene& scene, Film film) = 0;
};

struct PTRendererParams {
    int worker_count = 0;
    int task_tile_size = 16;
    int spp = 1;
    int min_depth = 3;
    int max_depth = 10;
    int direct_light_sample_num = 1;
};
RC<Renderer> create_pt_renderer(const PTRendererParams& params);

// GPU controls without a reference counterpart.
struct GpuRenderOptions {
    int device = 0;
    uint64_t seed = 0;
    int samples_per_pass = 0;        // 0 = automatic
    int shadow_transmittance = 0;    // 0 = closed form, 1 = ratio tracking
    int guide = 0;                   // 1 = learned path guiding, 2 = learned light sampling alone
    int guide_train_passes = 8;
    float guide_fraction = 0.5f;
    float guide_lr = 1e-3f;
    float guide_train_fraction = 0.0f;      // guide probability during training passes
    int guide_max_depth = 0;                // 0 = guide every depth
    int tail_paths = 0;                     // plain passes: one kernel finishes the last paths (0 = 131072, < 0 = off)
    int guide_light_grid = 0;               // learned light sampling: cells per axis (guide 1 or 2), 0 = off
    tr_render_stats* stats_out = nullptr;   // optional
};
RC<Renderer> create_pt_renderer(const PTRendererParams& params, const GpuRenderOptions& gpu);

// src/factory/renderer.hpp:39-48.  worker_count and task_tile_size are CPU
// scheduling controls without meaning on the GPU.
struct BDPTRendererParams {
    int worker_count = 0;
    int task_tile_size = 16;
    int max_camera_vertex_count = 10;
    int max_light_vertex_count = 10;
    int spp = 1;
};
RC<Renderer> create_bdpt_renderer(const BDPTRendererParams& params);
// GpuRenderOptions used: device, seed, samples_per_pass, stats_out.
RC<Renderer> create_bdpt_renderer(const BDPTRendererParams& params, const GpuRenderOptions& gpu);

}  // namespace tracer
