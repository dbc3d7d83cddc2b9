"""CPU oracle of the learned path-guiding network (test infrastructure only).

The guide has no counterpart in the reference renderer: this module restates,
in float64 numpy, the definitions the HIP kernels implement (input encoding,
MLP, von Mises-Fisher mixture head, training loss, backward pass, Adam), so
tests can compare the GPU network with it.  With ``bf16=True`` the forward and
backward passes round to bfloat16 at the points where the kernels hold bf16
data (layer inputs, weights, output and hidden deltas), accumulating in high
precision as the MFMA units accumulate in FP32.

How the oracle is pinned (tests/test_oracle_guide.py):
  * encoding: closed-form values at chosen points;
  * vMF lobe and mixture density: normalisation to 1 on the sphere by
    quadrature, and the closed-form constant kappa / (4 pi sinh kappa) e^kappa;
  * sampler: chi-square of sampled directions against the density;
  * loss gradient and the MLP backward pass: central finite differences of the
    float64 loss;
  * Adam: the update of a one-parameter quadratic by hand.
"""
from __future__ import annotations

import numpy as np

N_IN, N_IN_USED, N_OUT, N_LOBES, N_LAYERS = 96, 81, 48, 8, 5
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -4.0, 10.0
Q_EPS = 1e-3


# ---- bfloat16 rounding ---------------------------------------------------------
def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


# ---- layout --------------------------------------------------------------------
def perm_inverse(r: int) -> tuple[int, int]:
    """Lane group g and slot j holding feature r (0..31) of a 32-wide K step."""
    if r < 16:
        return r // 4, r % 4
    return (r - 16) // 4, 4 + (r - 16) % 4


def layer_dims(H: int) -> list[tuple[int, int]]:
    return [(N_IN, H), (H, H), (H, H), (H, H), (H, N_OUT)]


def unpack_params(p: np.ndarray, H: int):
    """Flat parameter vector -> [W_l (in x out)], [b_l]."""
    p = np.asarray(p, np.float64)
    Ws, bs, o = [], [], 0
    for fi, fo in layer_dims(H):
        Ws.append(p[o:o + fi * fo].reshape(fi, fo))
        o += fi * fo
    for _, fo in layer_dims(H):
        bs.append(p[o:o + fo])
        o += fo
    assert o == p.size
    return Ws, bs


def pack_params(Ws, bs) -> np.ndarray:
    return np.concatenate([w.ravel() for w in Ws] + [b.ravel() for b in bs])


def num_params(H: int) -> int:
    return sum(fi * fo + fo for fi, fo in layer_dims(H))


def input_used() -> np.ndarray:
    used = np.zeros(N_IN, bool)
    for i in range(N_IN):
        g, j = perm_inverse(i % 32)
        used[i] = g < 3 or j < 3
    return used


# ---- encoding ------------------------------------------------------------------
def guide_box(lo, hi):
    """Encoding box as the library builds it from scene bounds (float32)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    ext = hi - lo
    pad = np.float32(1e-3)
    l = np.where(ext > 0, lo - pad * ext, lo - pad)
    e = np.where(ext > 0, ext * (np.float32(1) + np.float32(2) * pad), np.float32(1))
    return l.astype(np.float32), (np.float32(1) / e).astype(np.float32)


def encode(x, n, wo, box_lo, inv_ext) -> np.ndarray:
    """Inputs (N, 96) of shading points: input i = 32 s + perm(g, j) is, for axis
    s, lane group g < 3: sin (j even) / cos (j odd) of 2 pi frac(2^k p_s) with
    k = 4 g + j // 2; lane group 3: p_s, n_s, wo_s for j = 0, 1, 2, else 0."""
    x = np.asarray(x, np.float32)
    p = np.clip((x - box_lo) * inv_ext, 0.0, 1.0).astype(np.float32).astype(np.float64)
    n = np.asarray(n, np.float64)
    wo = np.asarray(wo, np.float64)
    F = np.zeros((x.shape[0], N_IN))
    for i in range(N_IN):
        s, r = divmod(i, 32)
        g, j = perm_inverse(r)
        if g < 3:
            k = 4 * g + j // 2
            y = np.ldexp(p[:, s], k)
            fr = y - np.floor(y)
            F[:, i] = np.sin(2 * np.pi * fr) if j % 2 == 0 else np.cos(2 * np.pi * fr)
        elif j == 0:
            F[:, i] = p[:, s]
        elif j == 1:
            F[:, i] = n[:, s]
        elif j == 2:
            F[:, i] = wo[:, s]
    return F


# ---- MLP -----------------------------------------------------------------------
def forward(params, H, feats, bf16_emul=True):
    """Returns (raw (N, 48), layer inputs [a_0 .. a_4])."""
    Ws, bs = unpack_params(params, H)
    r = bf16 if bf16_emul else (lambda v: np.asarray(v, np.float64))
    a = r(feats)
    acts = [a]
    for l in range(4):
        z = a @ r(Ws[l]) + bs[l]
        a = r(np.maximum(z, 0.0))
        acts.append(a)
    raw = a @ r(Ws[4]) + bs[4]
    return raw, acts


# ---- vMF mixture ---------------------------------------------------------------
def mixture(raw):
    """raw (N, 48) -> weights (N, 8), unit means (N, 8, 3), kappa (N, 8)."""
    raw = np.asarray(raw, np.float64)
    a = raw[:, 0:40:5]
    w = np.exp(a - a.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    v = np.stack([raw[:, 1:40:5], raw[:, 2:40:5], raw[:, 3:40:5]], axis=-1)
    lv = np.linalg.norm(v, axis=-1)
    mu = np.where(lv[..., None] > 1e-12, v / np.maximum(lv, 1e-300)[..., None], np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(raw[:, 4:40:5], LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    return w, mu, kappa


def vmf_norm(kappa):
    """kappa / (4 pi sinh kappa) * e^kappa = kappa / (2 pi (1 - e^{-2 kappa}))."""
    kappa = np.asarray(kappa, np.float64)
    return kappa / (2 * np.pi * (-np.expm1(-2 * kappa)))


def lobe_pdf(mu, kappa, d):
    return vmf_norm(kappa) * np.exp(kappa * (np.einsum("...k,...k->...", mu, d) - 1.0))


def mixture_pdf(raw, d):
    """Density (N,) of unit directions d (N, 3)."""
    w, mu, kappa = mixture(raw)
    return np.sum(w * lobe_pdf(mu, kappa, np.asarray(d, np.float64)[:, None, :]), axis=1)


def onb(n):
    """Orthonormal tangents of unit vectors n (N, 3) (Duff et al. 2017)."""
    s = np.copysign(1.0, n[:, 2])
    a = -1.0 / (s + n[:, 2])
    b = n[:, 0] * n[:, 1] * a
    t = np.stack([1 + s * n[:, 0] ** 2 * a, s * b, -s * n[:, 0]], axis=1)
    bb = np.stack([b, s + n[:, 1] ** 2 * a, -n[:, 1]], axis=1)
    return t, bb


def mixture_sample(raw, u):
    """Directions (N, 3) from uniforms u (N, 3): lobe by weight, then the
    inverse-CDF vMF sampler of Jakob (2012)."""
    w, mu, kappa = mixture(raw)
    c = np.cumsum(w, axis=1)
    k = np.minimum((u[:, :1] >= c[:, :-1]).sum(axis=1), N_LOBES - 1)
    rows = np.arange(len(k))
    m, kap = mu[rows, k], kappa[rows, k]
    arg = np.maximum(u[:, 1] + (1 - u[:, 1]) * np.exp(-2 * kap), 1e-300)
    wz = np.clip(1 + np.log(arg) / kap, -1, 1)
    r = np.sqrt(np.maximum(0, 1 - wz ** 2))
    phi = 2 * np.pi * u[:, 2]
    t, b = onb(m)
    d = t * (r * np.cos(phi))[:, None] + b * (r * np.sin(phi))[:, None] + m * wz[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


# ---- loss ----------------------------------------------------------------------
def loss(raw, wi, weight):
    """Per-record loss -weight * log(q(wi) + 1e-3)."""
    return -np.asarray(weight, np.float64) * np.log(mixture_pdf(raw, wi) + Q_EPS)


def loss_grad(raw, wi, weight):
    """Analytic gradient (N, 48) of the per-record loss w.r.t. the head outputs."""
    raw = np.asarray(raw, np.float64)
    wi = np.asarray(wi, np.float64)
    weight = np.asarray(weight, np.float64)
    w, mu, kappa = mixture(raw)
    c = np.einsum("nkj,nj->nk", mu, wi)
    f = vmf_norm(kappa) * np.exp(kappa * (c - 1))
    q = np.sum(w * f, axis=1)
    dq = -weight / (q + Q_EPS)
    g = np.zeros_like(raw)
    pf = w * f
    g[:, 0:40:5] = dq[:, None] * w * (f - q[:, None])
    v = np.stack([raw[:, 1:40:5], raw[:, 2:40:5], raw[:, 3:40:5]], axis=-1)
    lv = np.linalg.norm(v, axis=-1)
    gm = (wi[:, None, :] - mu * c[..., None]) * (pf * kappa / np.maximum(lv, 1e-300))[..., None]
    gm = np.where(lv[..., None] > 1e-12, gm, 0.0)
    for a in range(3):
        g[:, 1 + a:40:5] = dq[:, None] * gm[..., a]
    s = raw[:, 4:40:5]
    with np.errstate(over="ignore"):   # 2 kappa / expm1(2 kappa) -> 0 for large kappa
        dlnf = 1 - 2 * kappa / np.expm1(2 * kappa) + kappa * (c - 1)
    inside = (s > LOG_KAPPA_MIN) & (s < LOG_KAPPA_MAX)
    g[:, 4:40:5] = np.where(inside, dq[:, None] * pf * dlnf, 0.0)
    return g


def normalized_weights(weight, clamp):
    """Training weights as rendering uses them: w * n / sum(w), clamped, / n."""
    w = np.asarray(weight, np.float64)
    n = w.size
    s = w[w > 0].sum()
    return np.minimum(w * n / s, clamp) / n


# ---- backward ------------------------------------------------------------------
def backward(params, H, feats, wi, weight, bf16_emul=True):
    """Summed loss and its gradient w.r.t. the flat parameters, following the
    kernels' rounding points when bf16_emul."""
    Ws, bs = unpack_params(params, H)
    r = bf16 if bf16_emul else (lambda v: np.asarray(v, np.float64))
    raw, acts = forward(params, H, feats, bf16_emul)
    weight = np.asarray(weight, np.float64)
    active = weight > 0
    L = float(np.sum(np.where(active, loss(raw, wi, np.where(active, weight, 0.0)), 0.0)))
    d = np.where(active[:, None], loss_grad(raw, wi, np.where(active, weight, 0.0)), 0.0)
    d = r(d)
    dW = [None] * 5
    db = [None] * 5
    dW[4] = acts[4].T @ d
    db[4] = d.sum(axis=0)
    g = d @ r(Ws[4]).T
    for l in range(3, -1, -1):
        d = r(g * (acts[l + 1] > 0))
        dW[l] = acts[l].T @ d
        db[l] = d.sum(axis=0)
        if l > 0:
            g = d @ r(Ws[l]).T
    return L, pack_params(dW, db)


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step (Kingma and Ba 2015) with bias correction; returns (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
