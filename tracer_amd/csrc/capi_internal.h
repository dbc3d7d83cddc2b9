// Context and helpers shared by the C ABI sources (capi.cpp, guide_capi.cpp).
#pragma once
#include "tracer_hip.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "envmap.h"
#include "tr_bdpt.h"
#include "tr_guide_kernels.h"
#include "tr_kernels.h"

namespace trcapi {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidArg : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Device allocation that grows on demand and never shrinks.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&ptr, n);
        if (e != hipSuccess) {
            ptr = nullptr;
            bytes = 0;
            (void)hipGetLastError();
            throw std::bad_alloc();
        }
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
};

struct HostTexture {
    int w = 0, h = 0;
    std::vector<uint8_t> rgb8;   // 8-bit texels, or
    std::vector<float> rgbf;     // float texels
};

struct HostScene {
    std::vector<float> P, N, UV;
    std::vector<HostTexture> tex;
    std::vector<int32_t> idx, mat, med_in, med_out;
    std::vector<float> emission;
    std::vector<tr_material> mats;
    std::vector<tr_medium> media;
    std::vector<std::vector<float>> grids;   // density grid per medium (empty: homogeneous)
    std::vector<std::vector<float>> albedo_grids, g_grids;   // per medium, empty: constant
    int has_env = 0, has_sun = 0;
    float env[3] = {0, 0, 0}, sun_dir[3] = {0, 0, 1}, sun_E[3] = {0, 0, 0};
    tr_camera cam{};
    float lo[3], hi[3];
    EnvMapHost envmap;   // w > 0: image environment
};

// Guide network, its optimiser state and the buffers of guided rendering.
struct GuideState {
    bool ready = false;
    tr_guide_config cfg{};
    bool box_set = false;       // box given by tr_guide_set_box (else the scene bounds)
    trhip::GuideBox box{};
    int step = 0;               // Adam steps taken
    DevBuf master, m, v, fwd, bwd;
    DevBuf raw;                 // head outputs per path slot
    DevBuf grec, rec, rec_aux;  // vertex records of training passes
    DevBuf sh_rec;              // light-sample record of each shadow ray
    DevBuf train;               // weighted training samples, in key order
    DevBuf train_raw;           // the same as the finalize kernel appended them
    DevBuf keys, vals;          // sort keys and sample indices, two halves each (in, sorted)
    DevBuf sort_tmp, wpart;     // radix-sort scratch, per-workgroup weight sums
    size_t sort_tmp_bytes = 0;
    DevBuf acts, deltas, pw, pb;
    DevBuf misc;                // [0] record count, [1] sample count (uint32), [2] weight sum, [3] loss (float)
    DevBuf grad;
    int64_t path_capacity = 0;
    int64_t record_capacity = 0;
    float avg_records_per_path = 0.0f;   // records per recording path; 0 = not measured yet
    double last_recording_paths = 0.0;
    int64_t last_samples = 0;   // weighted samples of the latest training pass (in `train`)
    trhip::GuideDims dims() const { return trhip::GuideDims{cfg.hidden}; }
    trhip::GuideNet net() const {
        trhip::GuideNet n;
        n.fwd = fwd.as<const trhip::bf16x8>();
        n.bwd = bwd.as<const trhip::bf16x8>();
        n.bias = master.as<const float>() + dims().num_weights();
        n.box = box;
        return n;
    }
};

// Learned light-sample distribution (tr_render_params::guide_light_grid):
// training sums, sampling tables and the training info of shadow rays.  The
// sums accumulate over every training pass until the scene, the grid or the
// guide is reset.
struct LightCacheState {
    int grid = 0, strata = 0, bins = 0, num_area = 0;
    float lo[3] = {0, 0, 0}, scale[3] = {0, 0, 0};
    DevBuf sum, hits, pmf, table, work, sh_g;
    int64_t sh_cap = 0;
    bool built = false;   // the tables hold a distribution from at least one training pass
};

// Bidirectional path tracing buffers (tr_render_bdpt), sized for `slots`
// camera samples with nc + nl vertices and K strategies each.
struct BdptState {
    DevBuf cv, lv, ncv, nlv, gray, ghit, gbeta, grng, gq0, gq1, con, sray, sidx, rad, splat, w0;
    int64_t slots = 0, vertices = 0, entries = 0, pixels = 0;
};

}  // namespace trcapi

struct tr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    trcapi::HostScene hs;
    bool have_scene = false, have_bvh = false;
    bool has_delta = false;   // a material samples specular (delta) lobes: paths may take extra bounces
    int film_w = 0, film_h = 0;   // size of the film of the last tr_render (kept on the device)
    // scene on device
    trcapi::DevBuf nodes, tris, tri_shade, tri_cls, mats, media, lights, light_cdf, light_pmf, light_guide;
    trcapi::DevBuf env_tex, env_cond_cdf, env_cond_func, env_marg_cdf, env_marg_func;
    trcapi::DevBuf tri_uv, tex_info, tex8, texf;   // material textures
    trcapi::DevBuf tri_tan;                        // normal derivatives (anisotropic materials)
    trcapi::DevBuf mat_tex, mat_raw;               // per-hit Disney parameters: texture ids, constants
    std::vector<int32_t> leaf_of;                  // leaf-order index of each primitive
    trcapi::DevBuf density;                        // voxels of every grid medium
    int64_t num_nodes = 0;
    int max_stack = 0;
    int bvh_depth = 0;
    trhip::DevScene ds{};
    // traversal spill area
    trcapi::DevBuf ovf_node, ovf_t;
    int ovf_threads = 0;
    // wavefront buffers
    trcapi::DevBuf ray_o, ray_d, hit, thr, rad, rng, info, q0, q1, sh_o, sh_d, sh_c, counters, stats, film;
    int64_t path_capacity = 0;
    // scratch for batch calls
    trcapi::DevBuf s_in, s_out, s_aux, s_aux2, s_aux3, s_aux4;
    std::vector<hipEvent_t> events;
    trcapi::GuideState guide;
    trcapi::LightCacheState lcache;
    trcapi::BdptState bdpt;
};

namespace trcapi {

inline tr_status fail(tr_ctx* ctx, tr_status code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

template <class F> tr_status guarded(tr_ctx* ctx, F&& f) {
    if (!ctx) return TR_ERR_INVALID;
    try {
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        f();
        ctx->err.clear();
        return TR_OK;
    } catch (const InvalidArg& e) {
        return fail(ctx, TR_ERR_INVALID, e.what());
    } catch (const StateError& e) {
        return fail(ctx, TR_ERR_STATE, e.what());
    } catch (const HipError& e) {
        return fail(ctx, TR_ERR_HIP, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, TR_ERR_OOM, "device or host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, TR_ERR_INVALID, e.what());
    }
}

inline void upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
    b.reserve(bytes < 16 ? 16 : bytes);
    if (bytes) hip_check(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
}

// guide_capi.cpp: guided rendering support for tr_render.
// Allocates the per-path guide buffers for up to P paths (initialising the
// network with defaults when needed) and sets the encoding box.
void guide_prepare(tr_ctx* c, int64_t P, bool training);
// Head outputs for the paths of queue `cur` (before k_shade).
void guide_infer(tr_ctx* c, const trhip::DevPaths& p, int cur, int num_paths);
// Fields of DevPaths for guided rendering.
void guide_fill_paths(tr_ctx* c, trhip::DevPaths& p, bool training);
// Record threshold for a training pass of num_paths paths.
unsigned int guide_record_threshold(tr_ctx* c, int64_t num_paths, int max_depth);
// Clears the record counter before a training pass.
void guide_begin_pass(tr_ctx* c);
// Turns the pass's records into weighted samples and takes the configured
// Adam steps; returns the number of samples trained on (synchronises).
int64_t guide_train_pass(tr_ctx* c, const trhip::DevPaths& p, uint64_t pass_seed, float lr);

}  // namespace trcapi
