// Host-side launchers of the guide kernels (k_guide.hip).
#pragma once
#include "tr_guide.h"
#include "tr_kernels.h"

namespace trhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Parameters of the guide MLP as the kernels read them.
struct GuideNet {
    const bf16x8* fwd;    // forward fragments (A = W^T per layer)
    const bf16x8* bwd;    // backward fragments (A = W, K padded to 32)
    const float* bias;    // FP32 biases: hidden layers l * H, output 4 H
    GuideBox box;
};

// Training batch: records are 16 floats (x, n, wo, wi, weight, 3 unused);
// sample j of the batch reads record (j * mul + add) mod n_total.
struct GuideBatch {
    const float4* rec;
    int n_total;
    unsigned long long mul, add;
    int n, n_pad;               // samples in the batch, rounded up to 32
    const float* wsum;          // sum of the weights of all n_total records (device)
    float clamp;                // normalised weight clamp
    int normalize;              // 1: weight * n_total / wsum, clamped, / n; 0: raw weights
};

// Scratch of one training step.
struct GuideScratch {
    unsigned short* acts;       // [act_rows][n_pad] bf16, feature-major
    unsigned short* deltas;     // [delta_rows][n_pad] bf16
    float* partial_w;           // [chunks][num_weights]
    float* partial_b;           // [chunks][delta_rows]
    float* loss;                // loss sum (device)
};

constexpr int kGuideChunk = 4096;   // samples per weight-gradient partial sum

// Scene box of the position encoding, from the scene bounds.
inline GuideBox guide_box(const float lo[3], const float hi[3]) {
    GuideBox b;
    const float pad = 1e-3f;
    float l[3], e[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = hi[a] - lo[a];
        l[a] = lo[a] - pad * (ext > 0.0f ? ext : 1.0f);
        e[a] = ext > 0.0f ? ext * (1.0f + 2.0f * pad) : 1.0f;
    }
    b.lo = mk(l[0], l[1], l[2]);
    b.inv_ext = mk(1.0f / e[0], 1.0f / e[1], 1.0f / e[2]);
    return b;
}

void launch_guide_infer(const GuideNet& net, int H, const DevScene& s, const DevPaths& p, int cur, int max_paths,
                        float* raw_out, hipStream_t st);
void launch_guide_forward_points(const GuideNet& net, int H, const float* pts, int n, float* raw_out, hipStream_t st);
void launch_guide_train(const GuideNet& net, int H, const GuideBatch& b, const GuideScratch& s, hipStream_t st);
void launch_guide_grads(const GuideDims& d, const GuideScratch& s, int n_pad, hipStream_t st);
// Sums the gradient partials; update != 0 also takes one Adam step.
void launch_guide_adam(const GuideDims& d, const GuideScratch& s, int n_pad, float* master, float* m, float* v,
                       unsigned short* fwd, unsigned short* bwd, float lr, float b1, float b2, float eps, int step,
                       int update, float* grad_out, hipStream_t st);
void launch_guide_pack(const GuideDims& d, const float* master, unsigned short* fwd, unsigned short* bwd, hipStream_t st);
// Records of a training pass -> compacted training batch with weights.
void launch_guide_finalize(const DevPaths& p, int num_records, float4* raw, unsigned long long* keys, unsigned int* vals,
                           unsigned int* train_count, hipStream_t st);
// Sorts the finalized samples by key, writes them in that order to train and
// their weight sum to wsum, both independent of the order in which
// workgroups appended them.  temp: guide_order_temp_bytes(n) bytes; partial:
// one float per 256 samples.
size_t guide_order_temp_bytes(int n);
hipError_t launch_guide_order(int n, unsigned long long* keys, unsigned long long* keys_sorted, unsigned int* vals,
                              unsigned int* vals_sorted, void* temp, size_t temp_bytes, const float4* raw, float4* train,
                              float* partial, float* wsum, hipStream_t st);
void launch_guide_mixture(const float* raw, const float* dirs, const float* u, int n, float* samples,
                          float* pdf_samples, float* pdf_dirs, hipStream_t st);
void launch_mfma_selftest(const float* A, const float* B, float* D, hipStream_t st);

}  // namespace trhip
