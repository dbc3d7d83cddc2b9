// Host code under AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitizers.py builds and runs it): the oracle
// (oracle/refcpu/ref_pt.cpp) and the library's host side (scene builders,
// OBJ / MTL ingest, PNG and Radiance HDR readers and writers,
// post-processing, the 4-wide BVH builder, the environment-map
// distribution), over the benchmark configurations at reduced detail,
// degenerate scenes, and corrupted or truncated files.  The sanitizers stop
// the program at the first error; a failed check exits with status 1.
//
// usage: sanitize_harness <scratch directory>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "envmap.h"
#include "tracer/scenes.h"
#include "tracer_hip.h"

extern "C" {
// oracle/refcpu/ref_pt.cpp
void* ref_scene_create(const tr_scene_desc* d, int max_leaf);
void ref_scene_destroy(void* h);
void ref_light_pmf(void* h, float* pmf_out);
void ref_light_pick(void* h, const float* u, int64_t n, int32_t* idx, float* pmf);
void ref_film_splat(int W, int H, int filter, float radius, float alpha, const float* pos, const float* L, int64_t n,
                    double* sums);
void ref_trace_closest_mt(void* hs, const float* rays, float* hits, int64_t n, int threads);
void ref_trace_closest_wt_bvh(void* hs, const float* rays, float* hits, int64_t n, int threads);
void ref_trace_any_wt(void* hs, const float* rays, int32_t* occ, int64_t n, int threads);
void ref_bsdf_eval(const tr_material* mats, const int32_t* ids, const float* wi, const float* wo, const float* ng,
                   int64_t n, int precision, int fixed, double* f, double* pdf);
void ref_bsdf_sample(const tr_material* mats, const int32_t* ids, const float* wo, const float* ng, const float* u,
                     int64_t n, int precision, int fixed, double* wi, double* f, double* pdf);
void ref_grid_density(const tr_medium* m, const float* pts, int64_t n, float* out);
void ref_grid_albedo_g(const tr_medium* m, const float* pts, int64_t n, float* albedo, float* g);
void* ref_env_create(int w, int h, const float* texels, const float* world_to_light);
void ref_env_destroy(void* e);
void ref_env_sample(void* e, const float* u, int64_t n, float* wi, float* pdf, float* L);
void ref_env_eval(void* e, const float* dirs, int64_t n, int quirks, float* L, float* pdf);
void ref_hg_sample(float g, const float* wo, const float* u, int64_t n, float* wi, float* pdf);
void ref_camera_rays(void* hs, const float* film_uv, const float* lens_uv, int64_t n, float* rays);
double ref_render(void* hs, const tr_render_params* p, int quirks, int nee, int threads, double* sums,
                  uint64_t* rays_out);
}

namespace {

int g_checks = 0;
void check(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) throw std::runtime_error("check failed: " + what);
}

std::mt19937 rng(12345);
float uni(float a = 0.0f, float b = 1.0f) { return std::uniform_real_distribution<float>(a, b)(rng); }

std::vector<float> random_rays(const tr_scene_desc& d, int n) {
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t v = 0; v < d.num_vertices; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], d.positions[3 * v + k]);
            hi[k] = std::fmax(hi[k], d.positions[3 * v + k]);
        }
    std::vector<float> r(size_t(n) * 8);
    for (int i = 0; i < n; ++i) {
        float* q = &r[size_t(i) * 8];
        for (int k = 0; k < 3; ++k) q[k] = uni(lo[k], hi[k]);
        float dx = uni(-1, 1), dy = uni(-1, 1), dz = uni(-1, 1);
        const float l = std::sqrt(dx * dx + dy * dy + dz * dz) + 1e-6f;
        q[3] = 0.0f;
        q[4] = dx / l; q[5] = dy / l; q[6] = dz / l;
        q[7] = i % 7 == 0 ? 0.5f : 1e30f;
    }
    return r;
}

// The oracle and the BVH builder over one flattened scene.
void exercise_scene(const tr_scene_desc& d, int max_leaf, const char* name) {
    void* h = ref_scene_create(&d, max_leaf);
    check(h != nullptr, std::string(name) + ": oracle scene");
    const int n = 2048;
    std::vector<float> rays = random_rays(d, n), hits(size_t(n) * 4), hmt(size_t(n) * 4);
    std::vector<int32_t> occ(n);
    ref_trace_closest_wt_bvh(h, rays.data(), hits.data(), n, 4);
    ref_trace_closest_mt(h, rays.data(), hmt.data(), n, 4);
    ref_trace_any_wt(h, rays.data(), occ.data(), n, 4);
    for (int i = 0; i < n; ++i) {
        int32_t prim;
        std::memcpy(&prim, &hits[size_t(i) * 4 + 1], 4);
        check((prim >= 0) == (occ[i] != 0), std::string(name) + ": closest and any hit agree");
    }
    std::vector<float> pmf(size_t(d.num_triangles) + 2);
    ref_light_pmf(h, pmf.data());
    std::vector<float> u(256);
    for (float& x : u) x = uni();
    std::vector<int32_t> idx(256);
    std::vector<float> pk(256);
    ref_light_pick(h, u.data(), 256, idx.data(), pk.data());
    std::vector<float> fu(64 * 2), lu(64 * 2), cr(64 * 8);
    for (float& x : fu) x = uni();
    for (float& x : lu) x = uni();
    ref_camera_rays(h, fu.data(), lu.data(), 64, cr.data());
    tr_render_params p{};
    p.width = 12; p.height = 10; p.spp = 2; p.min_depth = 2; p.max_depth = 5; p.direct_light_sample_num = 1;
    p.filter_type = TR_FILTER_GAUSSIAN; p.filter_radius = 1.5f; p.filter_alpha = 0.6f; p.seed = 3;
    std::vector<double> sums(size_t(p.width) * p.height * 4);
    uint64_t rays_out[2] = {0, 0};
    for (int quirks = 0; quirks < 2; ++quirks)
        for (int nee = 0; nee < 2; ++nee) {
            ref_render(h, &p, quirks, nee, 4, sums.data(), rays_out);
            for (double v : sums) check(std::isfinite(v), std::string(name) + ": finite film sums");
        }
    ref_scene_destroy(h);

    trhip::BvhBuildOptions o;
    o.max_leaf = max_leaf;
    o.threads = 4;
    const trhip::BvhBuildResult b = trhip::build_bvh4(d.positions, d.num_vertices, d.indices, d.num_triangles, o);
    check(!b.nodes.empty() && int64_t(b.tris.size()) == d.num_triangles, std::string(name) + ": 4-wide BVH");
    int64_t leaf_tris = 0;
    for (const trhip::Node4& nd : b.nodes)
        for (int c = 0; c < 4; ++c)
            if (nd.child[c] < 0 && nd.child[c] != -1) leaf_tris += (~nd.child[c]) & 15;
    check(leaf_tris == d.num_triangles, std::string(name) + ": every triangle in one leaf");
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    f << text;
    check(bool(f), "write " + path);
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void images(const std::string& dir) {
    const int W = 37, H = 23;
    std::vector<uint8_t> rgb8(size_t(W) * H * 3);
    std::vector<float> rgbf(size_t(W) * H * 3);
    for (size_t i = 0; i < rgb8.size(); ++i) {
        rgb8[i] = uint8_t(i * 7 % 256);
        rgbf[i] = float(i % 97) * 0.37f;
    }
    rgbf[5] = std::numeric_limits<float>::infinity();
    rgbf[6] = std::nanf("");
    rgbf[7] = 3.0e38f;
    const char* err = nullptr;
    const std::string png = dir + "/t.png", hdr = dir + "/t.hdr";
    check(trs_write_png(png.c_str(), W, H, rgb8.data(), &err) == 0, "write png");
    check(trs_write_hdr(hdr.c_str(), W, H, rgbf.data(), &err) == 0, "write hdr");
    int w = 0, h = 0;
    std::vector<uint8_t> back8(rgb8.size());
    check(trs_read_png(png.c_str(), &w, &h, back8.data(), int64_t(back8.size()), &err) == 0 && w == W && h == H,
          "read png");
    std::vector<float> backf(rgbf.size());
    check(trs_read_hdr(hdr.c_str(), &w, &h, backf.data(), int64_t(backf.size()), &err) == 0 && w == W && h == H,
          "read hdr");
    for (float v : backf) check(std::isfinite(v), "hdr texels finite");
    // too small a destination, truncated and corrupted files: errors, no reads or writes out of bounds
    check(trs_read_png(png.c_str(), &w, &h, back8.data(), 10, &err) != 0, "png capacity");
    const std::string p8 = read_file(png), h8 = read_file(hdr);
    for (size_t cut : {size_t(8), size_t(30), p8.size() / 2, p8.size() - 5}) {
        write_file(dir + "/cut.png", p8.substr(0, cut));
        check(trs_read_png((dir + "/cut.png").c_str(), &w, &h, back8.data(), int64_t(back8.size()), &err) != 0,
              "truncated png rejected");
    }
    std::string bad = p8;
    bad[40] ^= 0x55;
    write_file(dir + "/bad.png", bad);
    check(trs_read_png((dir + "/bad.png").c_str(), &w, &h, back8.data(), int64_t(back8.size()), &err) != 0,
          "png with a bad checksum rejected");
    for (size_t cut : {size_t(5), size_t(40), h8.size() / 2}) {
        write_file(dir + "/cut.hdr", h8.substr(0, cut));
        trs_read_hdr((dir + "/cut.hdr").c_str(), &w, &h, backf.data(), int64_t(backf.size()), &err);
    }
    std::vector<float> post(rgbf);
    for (float& v : post) if (!std::isfinite(v)) v = 0.0f;
    check(trs_postprocess(post.data(), W, H, 1.0f, 2.2f) == 0, "postprocess");
    check(trs_postprocess(post.data(), W, H, -1.0f, 0.0f) == 0, "postprocess, no-op");

    // environment-map distribution (host side of the image environment)
    trcapi::EnvMapHost e;
    e.w = W;
    e.h = H;
    e.rgb.assign(rgbf.begin(), rgbf.end());
    for (float& v : e.rgb) if (!std::isfinite(v) || v > 1e6f) v = 1.0f;
    trcapi::build_envmap(e);
    check(e.marg_int > 0.0f && int(e.cond_cdf.size()) == e.nv * (e.nu + 1), "environment distribution");
    double out[3];
    trcapi::env_lookup(e, -0.25, 1.75, out);
    void* re = ref_env_create(W, H, e.rgb.data(), e.rot);
    std::vector<float> u(2 * 512), wi(3 * 512), pdf(512), L(3 * 512);
    for (float& x : u) x = uni();
    ref_env_sample(re, u.data(), 512, wi.data(), pdf.data(), L.data());
    ref_env_eval(re, wi.data(), 512, 0, L.data(), pdf.data());
    ref_env_eval(re, wi.data(), 512, 1, L.data(), pdf.data());
    ref_env_destroy(re);
}

void obj_files(const std::string& dir) {
    const std::string mtl =
        "newmtl white\nKd 0.7 0.7 0.7\nnewmtl glossy\nKd 0.4 0.4 0.3\nKs 0.3 0.3 0.3\nNs 20\n"
        "newmtl glass\nKs 0.9 0.9 0.9\nNi 1.5\nnewmtl light\nKd 0.5 0.5 0.5\nKe 10 10 10\n"
        "newmtl tex\nKd 1 1 1\nmap_Kd t.png\n";
    std::string obj = "mtllib s.mtl\n";
    obj += "v -1 -1 1\nv 1 -1 1\nv 1 -1 -1\nv -1 -1 -1\nv -1 1 -1\nv 1 1 -1\nv 1 1 1\nv -1 1 1\n";
    obj += "v -0.2 0.99 -0.2\nv 0.2 0.99 -0.2\nv 0.2 0.99 0.2\nv -0.2 0.99 0.2\n";
    obj += "vt 0 0\nvt 2 0\nvt 2 2\nvt 0 2\nvn 0 1 0\nvn 0 -1 0\nvn 0 0 1\n";
    obj += "usemtl tex\nf 1/1/1 2/2/1 3/3/1 4/4/1\nusemtl white\nf 5//2 6//2 7//2 8//2\nf 4 3 6 5\n";
    obj += "usemtl glossy\nf 1 4 5 8\nusemtl glass\nf 2 7 6 3\nusemtl light\nf 9 10 11 12\n";
    write_file(dir + "/s.mtl", mtl);
    write_file(dir + "/room.obj", obj);
    tr_camera cam{};
    cam.pos[2] = 0.9f; cam.target[2] = -1.0f; cam.up[1] = 1.0f;
    cam.fov_y = 1.2f; cam.focal_distance = 1.0f; cam.aspect = 1.0f;
    const char* err = nullptr;
    for (int fog = 0; fog < 2; ++fog) {
        trs_scene* s = trs_load_obj((dir + "/room.obj").c_str(), &cam, fog, &err);
        check(s != nullptr, std::string("obj room: ") + (err ? err : ""));
        exercise_scene(*trs_desc(s), trs_max_leaf(s), "obj room");
        trs_destroy(s);
    }
    std::vector<float> dens(4 * 5 * 6, 0.3f);
    float l2w[12] = {2, 0, 0, -1, 0, 2, 0, -1, 0, 0, 2, -1}, alb[3] = {0.8f, 0.8f, 0.8f};
    std::vector<float> alb_grid(4 * 5 * 6), g_grid(4 * 5 * 6);
    for (float& x : alb_grid) x = uni(0, 1);
    for (float& x : g_grid) x = uni(-0.9f, 0.9f);
    for (int with_grids = 0; with_grids < 2; ++with_grids) {
        trs_scene* g = trs_load_obj_grid((dir + "/room.obj").c_str(), &cam, dens.data(), 4, 5, 6, l2w, alb, 0.3f,
                                         with_grids ? alb_grid.data() : nullptr, with_grids ? g_grid.data() : nullptr,
                                         4, &err);
        check(g != nullptr, "obj room with a density grid");
        exercise_scene(*trs_desc(g), trs_max_leaf(g), "grid room");
        const tr_medium& m = trs_desc(g)->media[trs_desc(g)->num_media - 1];
        std::vector<float> pts(3 * 300), dout(300), aout(900), gout(300);
        for (float& x : pts) x = uni(-3, 3);
        ref_grid_density(&m, pts.data(), 300, dout.data());
        ref_grid_albedo_g(&m, pts.data(), 300, aout.data(), gout.data());
        trs_destroy(g);
    }
    // malformed inputs: an error, never a crash
    const char* broken[] = {
        "v 0 0 0\nv 1 0 0\nf 1 2 3\n",                       // index past the vertices
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -9\n",             // negative index past the start
        "v 0 0 x\nf 1 2 3\n",                                // not a number
        "mtllib missing.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl nope\nf 1 2 3\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//9 2//9 3//9\n",     // normal index out of range
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/5 2/5 3/5\n",  // texture index out of range
        "v 0 0 0\nf 1 1\n",                                  // a face of two corners
        "",
    };
    int k = 0;
    for (const char* b : broken) {
        const std::string path = dir + "/broken" + std::to_string(k++) + ".obj";
        write_file(path, b);
        trs_scene* s = trs_load_obj(path.c_str(), &cam, 0, &err);
        if (s) trs_destroy(s);
    }
    check(trs_load_obj((dir + "/absent.obj").c_str(), &cam, 0, &err) == nullptr, "missing OBJ rejected");
}

void materials() {
    // every material kind over random directions, including grazing and
    // below-surface ones, and extreme parameters
    std::vector<tr_material> mats;
    for (int type : {TR_MAT_LAMBERT, TR_MAT_DISNEY, TR_MAT_PHONG, TR_MAT_GLASS})
        for (int v = 0; v < 4; ++v) {
            tr_material m{};
            m.type = type;
            for (int k = 0; k < 3; ++k) { m.base_color[k] = v == 0 ? 0.0f : uni(); m.specular_scale[k] = uni(); }
            m.metallic = v == 1 ? 1.0f : uni();
            m.roughness = v == 2 ? 0.0f : uni();
            m.transmission = v == 3 ? 1.0f : uni();
            m.transmission_roughness = v == 2 ? 0.0f : uni();
            m.ior = v == 3 ? 1.0f : 1.0f + uni(0.01f, 1.5f);
            m.specular_tint = uni(); m.anisotropic = v == 1 ? 1.0f : uni();
            m.sheen = uni(); m.sheen_tint = uni(); m.clearcoat = uni(); m.clearcoat_gloss = v == 2 ? 1.0f : uni();
            m.shininess = v == 0 ? 0.0f : uni(1, 1000);
            mats.push_back(m);
        }
    const int n = 4096;
    std::vector<int32_t> ids(n);
    std::vector<float> wi(3 * n), wo(3 * n), ng(3 * n), u(3 * n);
    for (int i = 0; i < n; ++i) {
        ids[i] = i % int(mats.size());
        for (int k = 0; k < 3; ++k) { wi[3 * i + k] = uni(-1, 1); wo[3 * i + k] = uni(-1, 1); u[3 * i + k] = uni(); }
        ng[3 * i] = uni(-0.3f, 0.3f); ng[3 * i + 1] = uni(-0.3f, 0.3f); ng[3 * i + 2] = 1.0f;
        if (i % 11 == 0) { wi[3 * i + 2] = 0.0f; wo[3 * i + 2] = 1e-8f; }   // grazing
    }
    std::vector<double> f(3 * n), pdf(n), dwi(3 * n);
    for (int precision = 0; precision < 2; ++precision)
        for (int fixed = 0; fixed < 2; ++fixed) {
            ref_bsdf_eval(mats.data(), ids.data(), wi.data(), wo.data(), ng.data(), n, precision, fixed, f.data(),
                          pdf.data());
            ref_bsdf_sample(mats.data(), ids.data(), wo.data(), ng.data(), u.data(), n, precision, fixed, dwi.data(),
                            f.data(), pdf.data());
        }
    std::vector<float> hw(3 * 512), hu(2 * 512), hwi(3 * 512), hp(512);
    for (float& x : hw) x = uni(-1, 1);
    for (float& x : hu) x = uni();
    for (float g : {-0.95f, 0.0f, 0.3f, 0.95f}) ref_hg_sample(g, hw.data(), hu.data(), 512, hwi.data(), hp.data());
}

void film() {
    const int W = 9, H = 7, n = 2000;
    std::vector<float> pos(2 * n), L(3 * n);
    for (int i = 0; i < n; ++i) {
        pos[2 * i] = uni(-1.0f, float(W) + 1.0f);
        pos[2 * i + 1] = uni(-1.0f, float(H) + 1.0f);
        for (int k = 0; k < 3; ++k) L[3 * i + k] = i % 13 == 0 ? std::nanf("") : uni();
    }
    std::vector<double> sums(size_t(W) * H * 4);
    for (float r : {0.5f, 1.0f, 2.5f, 16.0f}) {
        ref_film_splat(W, H, TR_FILTER_GAUSSIAN, r, 0.6f, pos.data(), L.data(), n, sums.data());
        ref_film_splat(W, H, TR_FILTER_BOX, r, 0.0f, pos.data(), L.data(), n, sums.data());
    }
}

void degenerate_scene() {
    // zero-area and duplicate triangles, an emitter of zero area, zero-length
    // and reversed vertex normals
    std::vector<float> P = {0, 0, 0, 1, 0, 0, 0, 1, 0, 2, 0, 0, 3, 0, 0, 5, 5, 5, 0, 0, 0, 1, 0, 0, 0, 1, 0};
    std::vector<float> N = {0, 0, 0, 0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1};
    std::vector<int32_t> I = {0, 1, 2, 1, 3, 4, 5, 5, 5, 6, 7, 8, 0, 1, 2};
    std::vector<int32_t> M(5, 0);
    std::vector<float> E = {0, 0, 0, 4, 4, 4, 0, 0, 0, 1, 1, 1, 0, 0, 0};
    tr_material mat{};
    mat.type = TR_MAT_DISNEY;
    mat.base_color[0] = mat.base_color[1] = mat.base_color[2] = 0.5f;
    mat.roughness = 0.0f;
    mat.ior = 1.5f;
    tr_scene_desc d{};
    d.num_vertices = 9; d.positions = P.data(); d.normals = N.data();
    d.num_triangles = 5; d.indices = I.data(); d.material_ids = M.data(); d.emission = E.data();
    d.num_materials = 1; d.materials = &mat;
    d.has_env = 1; d.env_radiance[0] = d.env_radiance[1] = d.env_radiance[2] = 0.2f;
    d.camera.pos[2] = 3.0f; d.camera.up[1] = 1.0f; d.camera.fov_y = 0.8f; d.camera.focal_distance = 1.0f;
    d.camera.aspect = 1.0f;
    exercise_scene(d, 2, "degenerate scene");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        const double detail[6] = {0, 1.0, 1.0 / 64, 1.0 / 12, 1.0 / 16, 1.0 / 2048};
        for (int c = 1; c <= 5; ++c) {
            const char* err = nullptr;
            trs_scene* s = trs_create(c, uint64_t(c - 1), detail[c], &err);
            check(s != nullptr, "configuration " + std::to_string(c));
            tr_render_params p{};
            trs_render_defaults(s, &p);
            check(p.width > 0 && p.spp > 0, "render defaults");
            exercise_scene(*trs_desc(s), trs_max_leaf(s), ("configuration " + std::to_string(c)).c_str());
            trs_destroy(s);
        }
        images(dir);
        obj_files(dir);
        materials();
        film();
        degenerate_scene();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "sanitize_harness: %s\n", e.what());
        return 1;
    }
    std::printf("ok: %d checks\n", g_checks);
    return 0;
}
