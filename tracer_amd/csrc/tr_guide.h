// Learned path guiding: input encoding, the von Mises-Fisher mixture head and
// the MFMA fragment layouts of the guide MLP.  No reference counterpart.
//
// Network: 96 input features (81 used) -> H -> H -> H -> H -> 48 (40 used),
// ReLU after the four hidden layers, H = 64 or 128.  Inputs of a shading point
// x with shading normal n and outgoing direction wo, with
// p = clamp((x - box_lo) / (box_hi - box_lo), 0, 1):
//   input i = 32 s + perm(g, j) (see below) for axis s = 0..2, lane group
//   g = 0..3, slot j = 0..7:
//     g < 3:  sin (j even) / cos (j odd) of 2 pi frac(2^k p_s), k = 4 g + j / 2
//     g = 3:  j = 0: p_s, j = 1: n_s, j = 2: wo_s, j > 2: 0
// i.e. per axis the 12 octaves of the frequency encoding plus the position,
// normal and direction components: 3 + 72 + 3 + 3 used inputs.
// Output, per lobe k = 0..7: logit a_k, mean m_k (3, normalised), log-concentration
// s_k: mixture weights softmax(a), kappa_k = exp(clamp(s_k, -4, 10)).
//
// MFMA layout (v_mfma_f32_16x16x32_bf16, one wavefront):
//   A (16 x 32): lane l holds row l % 16, eight K values given by slot j
//   B (32 x 16): lane l holds column l % 16, the same eight K values
//   D (16 x 16 fp32): lane l holds rows 4 (l / 16) + i, i = 0..3, column l % 16
// Layers are evaluated transposed, out^T = W^T act^T, so activations are B
// operands (K = feature, N = sample) and an output fragment becomes the next
// layer's B operand without data movement if slot j of lane group g = l / 16
// stands for feature perm(g, j) = j < 4 ? 4g + j : 16 + 4g + (j - 4) within a
// 32-feature K step (D tiles 2s and 2s + 1 feed K step s).  Weights are
// stored pre-permuted into per-lane 16-byte fragments with the same mapping.
#pragma once
#include "tr_device.h"

namespace trhip {

constexpr int kGuideIn = 96;        // padded input features
constexpr int kGuideInUsed = 81;    // used input features
constexpr int kGuideOut = 48;       // padded output
constexpr int kGuideLobes = 8;
constexpr int kGuideFreqs = 12;
constexpr int kGuideLayers = 5;

TR_HD int perm_feature(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }
// inverse of perm_feature for r in [0, 32): lane group and slot
TR_HD void perm_inverse(int r, int& g, int& j) {
    if (r < 16) { g = r >> 2; j = r & 3; }
    else { g = (r - 16) >> 2; j = 4 + ((r - 16) & 3); }
}

// Layer geometry for hidden width H.
struct GuideDims {
    int H;
    TR_HD int in(int l) const { return l == 0 ? kGuideIn : H; }
    TR_HD int out(int l) const { return l == kGuideLayers - 1 ? kGuideOut : H; }
    // out rounded up to a multiple of 32 (K of the backward pass)
    TR_HD int out_k(int l) const { return (out(l) + 31) / 32 * 32; }
    TR_HD int w_offset(int l) const {   // fp32 master weights [in][out], then biases
        int o = 0;
        for (int i = 0; i < l; ++i) o += in(i) * out(i);
        return o;
    }
    TR_HD int num_weights() const { return w_offset(kGuideLayers); }
    TR_HD int b_offset(int l) const {
        int o = num_weights();
        for (int i = 0; i < l; ++i) o += out(i);
        return o;
    }
    TR_HD int num_params() const { return b_offset(kGuideLayers); }
    // bf16 fragment buffers: forward (A = W^T) and backward (A = W, K padded)
    TR_HD int fwd_offset(int l) const {
        int o = 0;
        for (int i = 0; i < l; ++i) o += in(i) * out(i);
        return o;
    }
    TR_HD int bwd_offset(int l) const {
        int o = 0;
        for (int i = 0; i < l; ++i) o += in(i) * out_k(i);
        return o;
    }
    TR_HD int fwd_size() const { return fwd_offset(kGuideLayers); }
    TR_HD int bwd_size() const { return bwd_offset(kGuideLayers); }
    TR_HD int act_rows() const { return kGuideIn + 4 * H; }        // stored activations per sample
    TR_HD int delta_rows() const { return 4 * H + kGuideOut; }     // stored deltas per sample
    TR_HD int act_row(int l) const { return l == 0 ? 0 : kGuideIn + (l - 1) * H; }   // input of layer l
    TR_HD int delta_row(int l) const { return l * H; }                               // output delta of layer l
};

// Index of W[in][out] of layer l inside its forward fragment buffer.
TR_HD int fwd_frag_index(const GuideDims& d, int l, int in, int out) {
    const int ksteps = d.in(l) / 32;
    const int mt = out >> 4, s = in >> 5;
    int g, j;
    perm_inverse(in & 31, g, j);
    const int lane = (out & 15) + 16 * g;
    return d.fwd_offset(l) + ((mt * ksteps + s) * 64 + lane) * 8 + j;
}
// Index of W[in][out] of layer l inside its backward fragment buffer (K = out padded).
TR_HD int bwd_frag_index(const GuideDims& d, int l, int in, int out) {
    const int ksteps = d.out_k(l) / 32;
    const int mt = in >> 4, s = out >> 5;
    int g, j;
    perm_inverse(out & 31, g, j);
    const int lane = (in & 15) + 16 * g;
    return d.bwd_offset(l) + ((mt * ksteps + s) * 64 + lane) * 8 + j;
}

// Scene box for the position encoding.
struct GuideBox {
    f3 lo;
    f3 inv_ext;
};

// True for the 81 inputs that carry a feature (the others are zero).
TR_HD bool guide_input_used(int in) {
    int g, j;
    perm_inverse(in & 31, g, j);
    return g < 3 || j < 3;
}

// ---- vMF mixture -----------------------------------------------------------
constexpr float kLogKappaMin = -4.0f, kLogKappaMax = 10.0f;

struct Mixture {
    float w[kGuideLobes];      // softmax weights
    f3 mu[kGuideLobes];        // unit means
    float kappa[kGuideLobes];
    float norm[kGuideLobes];   // kappa / (2 pi (1 - exp(-2 kappa)))
};

// All 40 head outputs finite: the mixture is usable for sampling.
TR_HD bool mixture_raw_valid(const float* raw) {
    bool ok = true;
    for (int f = 0; f < 5 * kGuideLobes; ++f) ok = ok && isfinite(raw[f]);
    return ok;
}

TR_HD Mixture mixture_from_raw(const float* raw) {
    Mixture m;
    float mx = -1e30f;
    for (int k = 0; k < kGuideLobes; ++k) mx = fmaxf(mx, raw[5 * k]);
    float sum = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] = expf(raw[5 * k] - mx);
        sum += m.w[k];
    }
    for (int k = 0; k < kGuideLobes; ++k) {
        m.w[k] /= sum;
        const f3 v = mk(raw[5 * k + 1], raw[5 * k + 2], raw[5 * k + 3]);
        const float l = length(v);
        m.mu[k] = l > 1e-12f ? v * (1.0f / l) : mk(0.0f, 0.0f, 1.0f);
        const float s = fminf(fmaxf(raw[5 * k + 4], kLogKappaMin), kLogKappaMax);
        m.kappa[k] = expf(s);
        m.norm[k] = m.kappa[k] / (2.0f * kPi * (-expm1f(-2.0f * m.kappa[k])));
    }
    return m;
}

// vMF density kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)) of each lobe, summed.
TR_HD float mixture_pdf(const Mixture& m, f3 w) {
    float q = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) q += m.w[k] * m.norm[k] * expf(m.kappa[k] * (dot(m.mu[k], w) - 1.0f));
    return q;
}

// Sample a direction: lobe by weight (u0), then the vMF lobe (Jakob, "Numerically
// stable sampling of the von Mises Fisher distribution on S^2", 2012).
TR_HD f3 mixture_sample(const Mixture& m, float u0, float u1, float u2) {
    int k = 0;
    float c = 0.0f;
    for (; k < kGuideLobes - 1; ++k) {
        c += m.w[k];
        if (u0 < c) break;
    }
    const float kap = m.kappa[k];
    // inverse CDF of the cosine to the mean; the argument of the log is kept
    // above zero where exp(-2 kappa) underflows
    const float arg = fmaxf(u1 + (1.0f - u1) * expf(-2.0f * kap), 1e-30f);
    const float wz = fminf(fmaxf(1.0f + logf(arg) / kap, -1.0f), 1.0f);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
    float sp, cp;
    sincosf(2.0f * kPi * u2, &sp, &cp);
    f3 t, b;
    onb(m.mu[k], t, b);
    return normalize(t * (r * cp) + b * (r * sp) + m.mu[k] * wz);
}

// The mixture folded into the hemisphere around the unit vector n: a sample
// below the plane of n is mirrored about it, so none is wasted where a
// reflecting surface scatters nothing.  The folded density is q(w) + q(w'),
// w' the mirror image of w, above the plane and 0 below.
TR_HD f3 fold_to_hemisphere(f3 w, f3 n) {
    const float c = dot(w, n);
    return c < 0.0f ? w - n * (2.0f * c) : w;
}
TR_HD float mixture_pdf_folded(const Mixture& m, f3 w, f3 n) {
    const float c = dot(w, n);
    if (c < 0.0f) return 0.0f;
    return mixture_pdf(m, w) + mixture_pdf(m, w - n * (2.0f * c));
}

// Loss -weight * log(q + q_eps) of one training record and its gradient with
// respect to the head outputs.  Its expectation over directions sampled with
// pdf p and weight F / p is the cross entropy of the mixture against the
// target F, up to a constant: the KL-divergence objective of neural importance
// sampling (Mueller et al. 2019).  row holds the 48 head outputs and is
// overwritten with the gradient (zero beyond the 40 used outputs); lobes are
// streamed so only a few values stay live.  A non-finite loss yields a zero
// gradient and a zero return.
constexpr float kGuideQEps = 1e-3f;

TR_HD void lobe_params(const float* row, int k, f3& mu, float& lv, float& kappa, float& norm) {
    const f3 v = mk(row[5 * k + 1], row[5 * k + 2], row[5 * k + 3]);
    lv = length(v);
    mu = lv > 1e-12f ? v * (1.0f / lv) : mk(0.0f, 0.0f, 1.0f);
    kappa = expf(fminf(fmaxf(row[5 * k + 4], kLogKappaMin), kLogKappaMax));
    norm = kappa / (2.0f * kPi * (-expm1f(-2.0f * kappa)));
}

TR_HD float mixture_loss_grad_row(float* row, f3 w, float weight) {
    float mx = -3.0e38f;
    for (int k = 0; k < kGuideLobes; ++k) mx = fmaxf(mx, row[5 * k]);
    float se = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) se += expf(row[5 * k] - mx);
    float q = 0.0f;
    for (int k = 0; k < kGuideLobes; ++k) {
        f3 mu;
        float lv, kap, nrm;
        lobe_params(row, k, mu, lv, kap, nrm);
        q += (expf(row[5 * k] - mx) / se) * nrm * expf(kap * (dot(mu, w) - 1.0f));
    }
    const float qc = q + kGuideQEps;
    const float loss = -weight * logf(qc);
    if (!isfinite(loss)) {
        for (int f = 0; f < kGuideOut; ++f) row[f] = 0.0f;
        return 0.0f;
    }
    const float dq = -weight / qc;   // dL/dq
    for (int k = 0; k < kGuideLobes; ++k) {
        f3 mu;
        float lv, kap, nrm;
        lobe_params(row, k, mu, lv, kap, nrm);
        const float wk = expf(row[5 * k] - mx) / se;
        const float c = dot(mu, w);
        const float fk = nrm * expf(kap * (c - 1.0f));
        const float pf = wk * fk;
        const float s = row[5 * k + 4];
        // softmax: dq/da_k = w_k (f_k - q)
        row[5 * k] = dq * wk * (fk - q);
        // mean: dq/dm_k = w_k f_k kappa_k (w - (mu.w) mu) / |m_k|
        const f3 gm = lv > 1e-12f ? (w - mu * c) * (pf * kap / lv) : f3s(0.0f);
        row[5 * k + 1] = dq * gm.x;
        row[5 * k + 2] = dq * gm.y;
        row[5 * k + 3] = dq * gm.z;
        // log-concentration: d ln f / d ln kappa = 1 - 2 kappa / (e^{2 kappa} - 1) + kappa (c - 1)
        const float dlnf = 1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (c - 1.0f);
        row[5 * k + 4] = (s > kLogKappaMin && s < kLogKappaMax) ? dq * pf * dlnf : 0.0f;
    }
    for (int f = 5 * kGuideLobes; f < kGuideOut; ++f) row[f] = 0.0f;
    return loss;
}

}  // namespace trhip
