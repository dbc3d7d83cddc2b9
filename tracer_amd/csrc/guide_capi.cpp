// C ABI of the path-guiding network (tr_guide_* in include/tracer_hip.h) and
// the guide steps of tr_render: buffers, inference before shading, and the
// training pass that turns recorded path vertices into Adam steps.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "capi_internal.h"

using namespace trhip;
using namespace trcapi;

namespace {

// Initial lobe means: the eight cube diagonals, so the untrained mixture
// (all logits 0, kappa = 1) covers the sphere evenly.
constexpr float kInvSqrt3 = 0.57735026918962576f;
constexpr float kLobeDirs[kGuideLobes][3] = {
    {kInvSqrt3, kInvSqrt3, kInvSqrt3},   {-kInvSqrt3, kInvSqrt3, kInvSqrt3}, {kInvSqrt3, -kInvSqrt3, kInvSqrt3},
    {-kInvSqrt3, -kInvSqrt3, kInvSqrt3}, {kInvSqrt3, kInvSqrt3, -kInvSqrt3}, {-kInvSqrt3, kInvSqrt3, -kInvSqrt3},
    {kInvSqrt3, -kInvSqrt3, -kInvSqrt3}, {-kInvSqrt3, -kInvSqrt3, -kInvSqrt3}};

// Multiplier of the batch permutation j -> (j * mul + add) mod n; a prime above
// any record count, hence coprime to it.
constexpr unsigned long long kPermMul = 2654435761ull;

uint64_t splitmix(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float unit_uniform(uint64_t& s) { return float(splitmix(s) >> 40) * (1.0f / 16777216.0f); }

void check_config(const tr_guide_config& c) {
    if (c.hidden != 64 && c.hidden != 128) throw InvalidArg("guide hidden width must be 64 or 128");
    if (c.batch < 32 || c.batch > (1 << 22)) throw InvalidArg("guide batch must lie in [32, 2^22]");
    // the training kernel addresses its activation buffer with 32-bit byte offsets
    if (int64_t(GuideDims{c.hidden}.act_rows()) * ((c.batch + 31) / 32 * 32) * 2 >= (int64_t(1) << 31))
        throw InvalidArg("guide batch too large for this hidden width");
    if (c.steps_per_pass < 0 || c.steps_per_pass > 1024) throw InvalidArg("guide steps_per_pass must lie in [0, 1024]");
    if (c.record_capacity < 1024 || c.record_capacity > (1 << 26))
        throw InvalidArg("guide record_capacity must lie in [1024, 2^26]");
    if (!(c.lr > 0.0f) || !(c.beta1 >= 0.0f && c.beta1 < 1.0f) || !(c.beta2 >= 0.0f && c.beta2 < 1.0f) ||
        !(c.eps > 0.0f))
        throw InvalidArg("invalid Adam hyper-parameters");
    if (!(c.weight_clamp > 0.0f)) throw InvalidArg("guide weight_clamp must be positive");
}

// He-uniform hidden weights over the used inputs; output weights scaled by 0.1
// so the initial mixture stays close to its biases.
std::vector<float> initial_params(const tr_guide_config& cfg) {
    const GuideDims d{cfg.hidden};
    std::vector<float> p(size_t(d.num_params()), 0.0f);
    uint64_t s = cfg.seed;
    for (int l = 0; l < kGuideLayers; ++l) {
        const int fan_in = l == 0 ? kGuideInUsed : d.H;
        const float scale = std::sqrt(6.0f / float(fan_in)) * (l == kGuideLayers - 1 ? 0.1f : 1.0f);
        for (int in = 0; in < d.in(l); ++in)
            for (int out = 0; out < d.out(l); ++out) {
                const float u = unit_uniform(s);
                const bool used = (l > 0 || guide_input_used(in)) && (l < kGuideLayers - 1 || out < 5 * kGuideLobes);
                p[size_t(d.w_offset(l) + in * d.out(l) + out)] = used ? (2.0f * u - 1.0f) * scale : 0.0f;
            }
    }
    const int b4 = d.b_offset(kGuideLayers - 1);
    for (int k = 0; k < kGuideLobes; ++k)
        for (int a = 0; a < 3; ++a) p[size_t(b4 + 5 * k + 1 + a)] = kLobeDirs[k][a];
    return p;
}

// Largest training batch whose activations stay below 2^31 bytes.
int64_t g_max_batch(tr_ctx* c) {
    const int64_t rows = c->guide.dims().act_rows();
    return std::min<int64_t>(int64_t(1) << 22, ((int64_t(1) << 31) / (rows * 2) - 1) / 32 * 32);
}

void require_ready(tr_ctx* c) {
    if (!c->guide.ready) throw StateError("the guide network is not initialised (tr_guide_init)");
}

void upload_params(tr_ctx* c, const float* params) {
    GuideState& g = c->guide;
    const GuideDims d = g.dims();
    const size_t np = size_t(d.num_params());
    hipStream_t st = c->stream;
    hip_check(hipMemcpyAsync(g.master.ptr, params, np * 4, hipMemcpyHostToDevice, st), "H2D guide params");
    hip_check(hipMemsetAsync(g.m.ptr, 0, np * 4, st), "guide Adam clear");
    hip_check(hipMemsetAsync(g.v.ptr, 0, np * 4, st), "guide Adam clear");
    // padding of the fragment buffers (backward K beyond 48 outputs) must stay zero
    hip_check(hipMemsetAsync(g.fwd.ptr, 0, size_t(d.fwd_size()) * 2, st), "guide fragments clear");
    hip_check(hipMemsetAsync(g.bwd.ptr, 0, size_t(d.bwd_size()) * 2, st), "guide fragments clear");
    launch_guide_pack(d, g.master.as<float>(), g.fwd.as<unsigned short>(), g.bwd.as<unsigned short>(), st);
    hip_check(hipGetLastError(), "k_guide_pack launch");
    g.step = 0;
}

void init_guide(tr_ctx* c, const tr_guide_config& cfg) {
    check_config(cfg);
    GuideState& g = c->guide;
    g.cfg = cfg;
    const GuideDims d = g.dims();
    const size_t np = size_t(d.num_params());
    g.master.reserve(np * 4);
    g.m.reserve(np * 4);
    g.v.reserve(np * 4);
    g.grad.reserve(np * 4);
    g.fwd.reserve(size_t(d.fwd_size()) * 2);
    g.bwd.reserve(size_t(d.bwd_size()) * 2);
    g.misc.reserve(64);
    const std::vector<float> p = initial_params(cfg);
    upload_params(c, p.data());
    hip_check(hipStreamSynchronize(c->stream), "guide init sync");
    g.avg_records_per_path = 0.0f;   // not measured yet
    g.last_samples = 0;
    g.box_set = false;   // back to the scene bounds until tr_guide_set_box
    g.ready = true;
    c->lcache.grid = 0;  // the learned light distribution starts again with the network
    c->lcache.built = false;
}

void ensure_box(tr_ctx* c) {
    GuideState& g = c->guide;
    if (g.box_set) return;
    if (!c->have_scene) throw StateError("no encoding box: upload a scene or call tr_guide_set_box");
    g.box = guide_box(c->hs.lo, c->hs.hi);
}

// Training scratch for a batch of n_pad samples at the current width (the
// sizes depend on both, so every call re-checks them; reserve only grows).
void ensure_batch(GuideState& g, int n_pad) {
    const GuideDims d = g.dims();
    const size_t nchunks = size_t((n_pad + kGuideChunk - 1) / kGuideChunk);
    g.acts.reserve(size_t(d.act_rows()) * size_t(n_pad) * 2);
    g.deltas.reserve(size_t(d.delta_rows()) * size_t(n_pad) * 2);
    g.pw.reserve(nchunks * size_t(d.num_weights()) * 4);
    g.pb.reserve(nchunks * size_t(d.delta_rows()) * 4);
}

void require_bytes(const DevBuf& b, size_t n, const char* what) {
    if (b.bytes < n) throw StateError(std::string("guide buffer too small: ") + what);
}

// One training step on samples (j * mul + add) mod n_total, j < n, of rec.
void train_step(tr_ctx* c, const float4* rec, int n_total, int n, unsigned long long mul, unsigned long long add,
                int normalize, int update, float lr, float* grad_dev) {
    GuideState& g = c->guide;
    const GuideDims d = g.dims();
    const int n_pad = (n + 31) / 32 * 32;
    ensure_batch(g, n_pad);
    const size_t nchunks = size_t((n_pad + kGuideChunk - 1) / kGuideChunk);
    require_bytes(g.acts, size_t(d.act_rows()) * size_t(n_pad) * 2, "activations");
    require_bytes(g.deltas, size_t(d.delta_rows()) * size_t(n_pad) * 2, "deltas");
    require_bytes(g.pw, nchunks * size_t(d.num_weights()) * 4, "weight partials");
    require_bytes(g.pb, nchunks * size_t(d.delta_rows()) * 4, "bias partials");
    require_bytes(g.master, size_t(d.num_params()) * 4, "parameters");
    require_bytes(g.fwd, size_t(d.fwd_size()) * 2, "forward fragments");
    require_bytes(g.bwd, size_t(d.bwd_size()) * 2, "backward fragments");
    if (grad_dev) require_bytes(g.grad, size_t(d.num_params()) * 4, "gradient");
    if (n < 1 || n > n_total) throw InvalidArg("training batch larger than its record set");
    hipStream_t st = c->stream;
    float* misc = g.misc.as<float>();
    hip_check(hipMemsetAsync(misc + 3, 0, 4, st), "guide loss clear");
    GuideBatch b;
    b.rec = rec;
    b.n_total = n_total;
    b.mul = mul;
    b.add = add;
    b.n = n;
    b.n_pad = n_pad;
    b.wsum = misc + 2;
    b.clamp = g.cfg.weight_clamp;
    b.normalize = normalize;
    GuideScratch s;
    s.acts = g.acts.as<unsigned short>();
    s.deltas = g.deltas.as<unsigned short>();
    s.partial_w = g.pw.as<float>();
    s.partial_b = g.pb.as<float>();
    s.loss = misc + 3;
    launch_guide_train(g.net(), d.H, b, s, st);
    launch_guide_grads(d, s, n_pad, st);
    if (update) ++g.step;
    launch_guide_adam(d, s, n_pad, g.master.as<float>(), g.m.as<float>(), g.v.as<float>(), g.fwd.as<unsigned short>(),
                      g.bwd.as<unsigned short>(), lr, g.cfg.beta1, g.cfg.beta2, g.cfg.eps, std::max(g.step, 1), update,
                      grad_dev, st);
    hip_check(hipGetLastError(), "guide training launch");
}

}  // namespace

namespace trcapi {

// Record buffers (vertex records, samples, sort keys and scratch) for cap
// records; reserve only grows.
void ensure_records(tr_ctx* c, int64_t cap_records) {
    GuideState& g = c->guide;
    if (cap_records <= g.record_capacity) return;
    const size_t cap = size_t(cap_records);
    g.rec.reserve(cap * 64);
    g.rec_aux.reserve(cap * 32);
    g.train.reserve(cap * 64);
    g.train_raw.reserve(cap * 64);
    g.keys.reserve(cap * 16);
    g.vals.reserve(cap * 8);
    g.wpart.reserve((cap / 256 + 1) * 4);
    g.sort_tmp_bytes = guide_order_temp_bytes(int(cap));
    if (g.sort_tmp_bytes == 0) throw HipError("radix sort scratch query failed");
    g.sort_tmp.reserve(g.sort_tmp_bytes);
    g.record_capacity = cap_records;
}

void guide_prepare(tr_ctx* c, int64_t P, bool training) {
    GuideState& g = c->guide;
    if (!g.ready) {
        tr_guide_config cfg;
        tr_guide_default_config(&cfg);
        init_guide(c, cfg);
    }
    ensure_box(c);
    if (P > g.path_capacity) {
        g.raw.reserve(size_t(P) * kGuideOut * 4);
        g.grec.reserve(size_t(P) * 4);
        g.sh_rec.reserve(size_t(P) * 4);   // one shadow ray per path and bounce at most
        g.path_capacity = P;
    }
    if (training) ensure_records(c, g.cfg.record_capacity);
}

void guide_fill_paths(tr_ctx* c, DevPaths& p, bool training) {
    GuideState& g = c->guide;
    p.graw = g.raw.as<float>();
    if (training) {
        p.grec = g.grec.as<int>();
        p.rec = g.rec.as<float4>();
        p.rec_aux = g.rec_aux.as<float4>();
        p.rec_count = g.misc.as<unsigned int>();
        p.rec_cap = int(g.record_capacity);
        p.sh_rec = g.sh_rec.as<int>();
    }
}

void guide_infer(tr_ctx* c, const DevPaths& p, int cur, int num_paths) {
    GuideState& g = c->guide;
    launch_guide_infer(g.net(), g.cfg.hidden, c->ds, p, cur, num_paths, g.raw.as<float>(), c->stream);
}

// Share of the paths that record: about record_capacity (the config's
// records per pass) at the records per path the previous pass measured.  The
// buffers then grow to hold every record the recording paths can make (a
// continuation and a light-sample record per surface vertex), so none is
// dropped: which records a full buffer drops would depend on the order in
// which workgroups append them, and training would not be reproducible.
unsigned int guide_record_threshold(tr_ctx* c, int64_t num_paths, int max_depth) {
    GuideState& g = c->guide;
    const double per_path = g.avg_records_per_path > 0.0f ? double(g.avg_records_per_path) : 12.0;
    const double want = double(g.cfg.record_capacity) / (double(num_paths) * per_path);
    const unsigned int thr = want >= 1.0 ? 0xFFFFFFFFu : unsigned(want * 4294967296.0);
    g.last_recording_paths = double(num_paths) * (double(thr) / 4294967296.0);
    // the slot hash picks close to this many paths: 5% and 65536 paths of margin
    const double most = (g.last_recording_paths * 1.05 + 65536.0) * 2.0 * double(max_depth + 1);
    ensure_records(c, int64_t(std::min(most, double(int64_t(1) << 26))));
    return thr;
}

void guide_begin_pass(tr_ctx* c) {
    hip_check(hipMemsetAsync(c->guide.misc.ptr, 0, 16, c->stream), "guide counters clear");
}

int64_t guide_train_pass(tr_ctx* c, const DevPaths& p, uint64_t pass_seed, float lr) {
    GuideState& g = c->guide;
    hipStream_t st = c->stream;
    unsigned int counts[2] = {0, 0};
    hip_check(hipMemcpyAsync(counts, g.misc.ptr, 4, hipMemcpyDeviceToHost, st), "D2H record count");
    hip_check(hipStreamSynchronize(st), "guide sync");
    const int nrec = int(std::min<int64_t>(counts[0], g.record_capacity));
    unsigned long long* keys = g.keys.as<unsigned long long>();
    unsigned int* vals = g.vals.as<unsigned int>();
    const size_t cap = size_t(g.record_capacity);
    launch_guide_finalize(p, nrec, g.train_raw.as<float4>(), keys, vals, g.misc.as<unsigned int>() + 1, st);
    hip_check(hipMemcpyAsync(counts, g.misc.ptr, 8, hipMemcpyDeviceToHost, st), "D2H sample count");
    hip_check(hipStreamSynchronize(st), "guide sync");
    // samples in key order and their weight sum, independent of the order in
    // which workgroups appended them: training is reproducible
    hip_check(launch_guide_order(int(counts[1]), keys, keys + cap, vals, vals + cap, g.sort_tmp.ptr, g.sort_tmp_bytes,
                                 g.train_raw.as<float4>(), g.train.as<float4>(), g.wpart.as<float>(), g.misc.as<float>() + 2,
                                 st),
              "guide sample sort");
    // records per recording path (the counter keeps counting past the
    // capacity), for the next pass's threshold
    if (counts[0] > 0 && g.last_recording_paths > 0.0)
        g.avg_records_per_path = std::max(0.05f, float(double(counts[0]) / g.last_recording_paths));
    const int ntrain = int(counts[1]);
    g.last_samples = ntrain;
    if (ntrain < 32 || g.cfg.steps_per_pass == 0) return 0;
    const int n = std::min(g.cfg.batch, ntrain);
    uint64_t s = pass_seed ^ 0x6775696465ull;
    for (int k = 0; k < g.cfg.steps_per_pass; ++k) {
        const unsigned long long add = splitmix(s) % (unsigned long long)ntrain;
        train_step(c, g.train.as<float4>(), ntrain, n, kPermMul % (unsigned long long)ntrain, add, 1, 1, lr, nullptr);
    }
    return int64_t(n) * g.cfg.steps_per_pass;
}

}  // namespace trcapi

extern "C" {

void tr_guide_default_config(tr_guide_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->hidden = 64;
    cfg->batch = 1 << 18;
    cfg->steps_per_pass = 16;
    cfg->record_capacity = 1 << 20;
    cfg->lr = 1e-3f;
    cfg->beta1 = 0.9f;
    cfg->beta2 = 0.99f;
    cfg->eps = 1e-8f;
    cfg->weight_clamp = 1000.0f;
    cfg->seed = 0x6775696465ull;
}

tr_status tr_guide_init(tr_ctx* c, const tr_guide_config* cfg) {
    return guarded(c, [&] {
        tr_guide_config k;
        if (cfg) k = *cfg;
        else tr_guide_default_config(&k);
        init_guide(c, k);
    });
}

tr_status tr_guide_num_params(tr_ctx* c, int64_t* n) {
    return guarded(c, [&] {
        require_ready(c);
        if (!n) throw InvalidArg("n is NULL");
        *n = c->guide.dims().num_params();
    });
}

tr_status tr_guide_get_params(tr_ctx* c, float* params, int64_t n) {
    return guarded(c, [&] {
        require_ready(c);
        if (!params || n != c->guide.dims().num_params()) throw InvalidArg("params must hold num_params floats");
        hip_check(hipMemcpyAsync(params, c->guide.master.ptr, size_t(n) * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_guide_set_params(tr_ctx* c, const float* params, int64_t n) {
    return guarded(c, [&] {
        require_ready(c);
        if (!params || n != c->guide.dims().num_params()) throw InvalidArg("params must hold num_params floats");
        upload_params(c, params);
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_guide_set_box(tr_ctx* c, const float lo[3], const float hi[3]) {
    return guarded(c, [&] {
        if (!lo || !hi) throw InvalidArg("box is NULL");
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] >= lo[a])) throw InvalidArg("invalid box");
        c->guide.box = guide_box(lo, hi);
        c->guide.box_set = true;
    });
}

tr_status tr_guide_forward(tr_ctx* c, const float* points, int64_t n, float* raw) {
    return guarded(c, [&] {
        require_ready(c);
        ensure_box(c);
        if (n < 0 || n > (int64_t(1) << 26)) throw InvalidArg("n must lie in [0, 2^26]");
        if (n == 0) return;
        if (!points || !raw) throw InvalidArg("NULL array");
        hipStream_t st = c->stream;
        upload(c->s_in, points, size_t(n) * 36, st);
        c->s_out.reserve(size_t(n) * kGuideOut * 4);
        launch_guide_forward_points(c->guide.net(), c->guide.cfg.hidden, c->s_in.as<float>(), int(n), c->s_out.as<float>(),
                                    st);
        hip_check(hipGetLastError(), "k_guide_forward_points launch");
        hip_check(hipMemcpyAsync(raw, c->s_out.ptr, size_t(n) * kGuideOut * 4, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

tr_status tr_guide_train_batch(tr_ctx* c, const float* records, int64_t n, int32_t flags, float* loss, float* grad) {
    return guarded(c, [&] {
        require_ready(c);
        ensure_box(c);
        if (n < 1 || n > g_max_batch(c)) throw InvalidArg("n must lie in [1, the largest batch of this width]");
        if (!records) throw InvalidArg("records is NULL");
        GuideState& g = c->guide;
        hipStream_t st = c->stream;
        upload(c->s_in, records, size_t(n) * 64, st);
        float misc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        double ws = 0.0;
        for (int64_t i = 0; i < n; ++i) ws += records[16 * i + 12] > 0.0f ? records[16 * i + 12] : 0.0f;
        misc[2] = float(ws);
        hip_check(hipMemcpyAsync(g.misc.ptr, misc, 16, hipMemcpyHostToDevice, st), "H2D");
        const int normalize = (flags & 2) ? 1 : 0, update = (flags & 1) ? 1 : 0;
        train_step(c, c->s_in.as<float4>(), int(n), int(n), 1ull, 0ull, normalize, update, g.cfg.lr, g.grad.as<float>());
        if (grad)
            hip_check(hipMemcpyAsync(grad, g.grad.ptr, size_t(g.dims().num_params()) * 4, hipMemcpyDeviceToHost, st),
                      "D2H grad");
        hip_check(hipMemcpyAsync(misc, g.misc.ptr, 16, hipMemcpyDeviceToHost, st), "D2H loss");
        hip_check(hipStreamSynchronize(st), "sync");
        if (loss) *loss = misc[3];
    });
}

tr_status tr_guide_mixture_eval(tr_ctx* c, const float* raw, const float* dirs, const float* u, int64_t n,
                                float* samples, float* pdf_samples, float* pdf_dirs) {
    return guarded(c, [&] {
        if (n < 0 || n > (int64_t(1) << 24)) throw InvalidArg("n must lie in [0, 2^24]");
        if (n == 0) return;
        if (!raw || !dirs || !u || !samples || !pdf_samples || !pdf_dirs) throw InvalidArg("NULL array");
        hipStream_t st = c->stream;
        const size_t nr = size_t(n) * kGuideOut * 4, n3 = size_t(n) * 12;
        c->s_in.reserve(nr + 2 * n3);
        char* in = c->s_in.as<char>();
        hip_check(hipMemcpyAsync(in, raw, nr, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(in + nr, dirs, n3, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(in + nr + n3, u, n3, hipMemcpyHostToDevice, st), "H2D");
        c->s_out.reserve(n3 + size_t(n) * 8);
        char* out = c->s_out.as<char>();
        launch_guide_mixture(reinterpret_cast<float*>(in), reinterpret_cast<float*>(in + nr),
                             reinterpret_cast<float*>(in + nr + n3), int(n), reinterpret_cast<float*>(out),
                             reinterpret_cast<float*>(out + n3), reinterpret_cast<float*>(out + n3 + size_t(n) * 4), st);
        hip_check(hipGetLastError(), "k_guide_mixture launch");
        hip_check(hipMemcpyAsync(samples, out, n3, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipMemcpyAsync(pdf_samples, out + n3, size_t(n) * 4, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipMemcpyAsync(pdf_dirs, out + n3 + size_t(n) * 4, size_t(n) * 4, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

tr_status tr_guide_last_samples(tr_ctx* c, float* out, int64_t cap, int64_t* n) {
    return guarded(c, [&] {
        require_ready(c);
        if (!n) throw InvalidArg("n is NULL");
        *n = c->guide.last_samples;
        const int64_t k = std::min<int64_t>(cap, c->guide.last_samples);
        if (k > 0) {
            if (!out) throw InvalidArg("out is NULL");
            hip_check(hipMemcpyAsync(out, c->guide.train.ptr, size_t(k) * 64, hipMemcpyDeviceToHost, c->stream), "D2H");
        }
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

tr_status tr_guide_mfma_selftest(tr_ctx* c, const float* A, const float* B, float* D) {
    return guarded(c, [&] {
        if (!A || !B || !D) throw InvalidArg("NULL array");
        hipStream_t st = c->stream;
        float host[16 * 32 + 32 * 16];
        std::memcpy(host, A, 16 * 32 * 4);
        std::memcpy(host + 16 * 32, B, 32 * 16 * 4);
        upload(c->s_in, host, sizeof host, st);
        c->s_out.reserve(16 * 16 * 4);
        launch_mfma_selftest(c->s_in.as<float>(), c->s_in.as<float>() + 16 * 32, c->s_out.as<float>(), st);
        hip_check(hipGetLastError(), "k_mfma_selftest launch");
        hip_check(hipMemcpyAsync(D, c->s_out.ptr, 16 * 16 * 4, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "sync");
    });
}

}  // extern "C"
