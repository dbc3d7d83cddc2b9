// Internal classes behind the tracer:: factory API and the scene flattener
// that turns a tracer::Scene into a tr_scene_desc for libtracer_hip.so.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "tracer/tracer.hpp"
#include "tracer_hip.h"

namespace tracer {
namespace detail {

class ConstantTexture2D : public Texture2D {
public:
    explicit ConstantTexture2D(const Spectrum& c) : c(c) {}
    Spectrum evaluate(const Point2f&) const override { return c; }
    bool is_constant() const override { return true; }
    Spectrum c;
};

// An image loaded by create_texture2d_from_file / create_texture_from_file:
// h rows of w RGB texels, float (rgb) or 8-bit (rgb8, read as value / 255),
// sampled bilinearly with wrap_repeat.
class ImageTexture2D : public Texture2D {
public:
    Spectrum evaluate(const Point2f& uv) const override;
    int w = 0, h = 0;
    std::vector<float> rgb;
    std::vector<uint8_t> rgb8;
    bool is8() const { return !rgb8.empty(); }
};
// Radiance RGBE (.hdr) reader; throws std::runtime_error on a read or format error.
void read_hdr(const std::string& path, int& width, int& height, std::vector<float>& rgb);
// 8-bit PNG reader (grey, RGB, palette, RGBA; grey + alpha is refused as the
// reference refuses 2-channel images), rows bottom to top as the reference's
// stbi_set_flip_vertically_on_load(true) returns them.
void read_png(const std::string& path, int& width, int& height, std::vector<uint8_t>& rgb);
// A texture named in an MTL material, scaled by its colour (image_file.cpp:100-127).
RC<Texture2D> texture_from_file_scaled(const std::string& name, const real* scale);

class DisneyMaterial : public Material {
public:
    RC<const Texture2D> base_color, metallic, roughness, transmission, transmission_roughness, ior, specular_scale,
        specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss;
    RC<const Texture2D> normal_map;   // NormalMapper's map, or null
};
class PhongMaterial : public Material {
public:
    RC<const Texture2D> ka, kd, ks, ns;
};
class GlassMaterial : public Material {
public:
    RC<const Texture2D> kr, kt, ior;
};

// triangle.cpp:15-33 — world-space vertices of one mesh
struct TriangleMesh {
    std::vector<Point3f> p;
    std::vector<Normal3f> n;
    std::vector<Point2f> uv;
    std::vector<int> indices;
};
class TriangleShape : public Shape {
public:
    TriangleShape(RC<const TriangleMesh> m, int i) : mesh(std::move(m)), index(i) {}
    RC<const TriangleMesh> mesh;
    int index;
};

class HomogeneousMedium : public Medium {
public:
    Spectrum sigma_a, sigma_s;
    real g = 0;
    int max_scattering_count = 0;
};
class VacuumMedium : public Medium {};
class HeterogeneousMedium : public Medium {
public:
    Transform local_to_world;
    RC<const Texture3D> density, albedo, g;
    int max_scattering_count = 0;
    bool white_for_indirect = false;
};
// Texture3D over a float volume (image_texture3d.cpp), or a constant
class ImageTexture3D : public Texture3D {
public:
    RC<const Image3D<real>> image;
    real max_value = 0;
    Spectrum evaluate(const Point3f& uvw) const override;
    real max_real() const override { return max_value; }
};
class ConstantTexture3D : public Texture3D {
public:
    Spectrum value;
    Spectrum evaluate(const Point3f&) const override { return value; }
    real max_real() const override { return value.r; }
    bool is_constant() const override { return true; }
};

class DiffuseLight : public AreaLight {
public:
    Spectrum emission;
};
class IBLConstant : public EnvironmentLight {
public:
    RC<const Texture2D> tex;
    real world_to_light[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};
class DirectionalLight : public Light {
public:
    Vector3f direction;
    Spectrum irradiance;
};

class GeometricPrimitive : public Primitive {
public:
    RC<Shape> shape;
    RC<Material> material;
    MediumInterface mi;
    Spectrum emission;
    Box<DiffuseLight> light;
    const AreaLight* as_area_light() const noexcept override { return light.get(); }
};

// Process-wide counter for the version stamps below.
uint64_t next_version();

class BVHAccel : public Aggregate {
public:
    explicit BVHAccel(int max_leaf) : max_leaf(max_leaf) {}
    void build(std::vector<RC<Primitive>> p) override {
        prims = std::move(p);
        version = next_version();
    }
    int max_leaf;
    std::vector<RC<Primitive>> prims;
    uint64_t version = 0;   // changes with every build: renderers re-upload the geometry
};

class ThinLensCamera : public Camera {
public:
    real aspect = 1, fov = 0, lens_radius = 0, focal_dist = 1;
    Point3f pos, target;
    Vector3f up;
};

class GaussianFilter : public Filter {
public:
    real r = 0.5f, alpha = 2.0f;
    real radius() const noexcept override { return r; }
};
class BoxFilter : public Filter {
public:
    real r = 0.5f;
    real radius() const noexcept override { return r; }
};

class GeneralScene : public Scene {
public:
    explicit GeneralScene(RC<Aggregate> a) : accel(std::move(a)), uid(next_version()) {}
    void set_camera(RC<const Camera> c) override { camera = std::move(c); }
    RC<const Camera> get_camera() const noexcept override { return camera; }
    void prepare_to_render() override {}
    RC<Aggregate> accel;
    RC<const Camera> camera;
    const uint64_t uid;   // distinct for every scene object, also one built at a freed one's address
};

// A scene flattened to plain arrays; desc() points into them.
struct FlatScene {
    std::vector<float> P, N, UV;
    std::vector<int32_t> idx, mat, med_in, med_out;
    std::vector<float> emission;
    std::vector<tr_material> mats;
    std::vector<tr_medium> media;
    std::vector<std::vector<float>> grids;   // voxels of each heterogeneous medium, media point here
    int has_env = 0, has_sun = 0;
    float env[3] = {0, 0, 0}, sun_dir[3] = {0, 0, 1}, sun_E[3] = {0, 0, 0};
    int env_w = 0, env_h = 0;             // image environment
    std::vector<float> env_pix;
    float env_rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::vector<ImageTexture2D> tex;   // copies of the material textures
    mutable std::vector<tr_texture> tex_desc;
    tr_camera cam{};
    int max_leaf = 3;
    tr_scene_desc desc() const;
};

// Throws std::runtime_error for scene content the GPU path does not support.
FlatScene flatten(const Scene& scene);
// The scene's camera alone (a ThinLensCamera, else std::runtime_error).
tr_camera camera_desc(const Scene& scene);

}  // namespace detail
}  // namespace tracer
