"""Pins of the guide oracle (oracle/guide_ref.py) by closed forms, quadrature,
chi-square and finite differences, so the GPU network is compared with a
checked reference.  The guide has no counterpart in the reference renderer."""
import numpy as np
import pytest
from scipy import stats

from oracle import guide_ref as G


def _rand_raw(rng, n, logk=(-1.0, 4.0)):
    raw = rng.normal(0.0, 1.0, (n, 48))
    raw[:, 4:40:5] = rng.uniform(*logk, (n, 8))
    raw[:, 40:] = 0.0
    return raw


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 20000), rng.uniform(-1e-3, 1e-3, 2000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 65504.0])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(G.bf16(x), ref)
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert G.bf16(np.float32(1.0 + 2 ** -8)) == 1.0


def test_encoding_closed_form():
    lo, inv = np.zeros(3, np.float32), np.ones(3, np.float32)
    x = np.array([[0.25, 0.5, 0.0]], np.float32)
    n = np.array([[0.0, 0.6, 0.8]])
    wo = np.array([[1.0, 0.0, 0.0]])
    F = G.encode(x, n, wo, lo, inv)[0]
    for s, p in enumerate([0.25, 0.5, 0.0]):
        for i in range(32):
            g, j = G.perm_inverse(i)
            v = F[32 * s + i]
            if g < 3:
                k = 4 * g + j // 2
                ang = 2 * np.pi * ((p * 2 ** k) % 1.0)
                assert v == pytest.approx(np.sin(ang) if j % 2 == 0 else np.cos(ang), abs=1e-12)
            else:
                assert v == pytest.approx([p, n[0, s], wo[0, s], 0, 0, 0, 0, 0][j], abs=1e-12)
    # p = 0.25: octave 0 gives (sin, cos) = (1, 0), octave 1 (0, -1), then (0, 1)
    assert F[0] == pytest.approx(1.0) and F[1] == pytest.approx(0.0, abs=1e-12)
    assert F[2] == pytest.approx(0.0, abs=1e-12) and F[3] == pytest.approx(-1.0)
    assert G.input_used().sum() == 81


def test_box_clamps_outside_points():
    lo, inv = G.guide_box([0, 0, 0], [2, 4, 8])
    F = G.encode(np.array([[-5.0, 100.0, 4.0]], np.float32), np.zeros((1, 3)), np.zeros((1, 3)), lo, inv)[0]
    g3 = [32 * s + 12 for s in range(3)]   # lane group 3, slot 0 (perm(3, 0) = 12): the position
    assert F[g3[0]] == 0.0 and F[g3[1]] == 1.0 and F[g3[2]] == pytest.approx(0.5, abs=1e-5)


@pytest.mark.parametrize("kappa", [np.exp(-4.0), 0.5, 1.0, 10.0, 300.0, np.exp(8.0), np.exp(10.0)])
def test_vmf_normalisation(kappa):
    # the density depends on t = mu . w only: integral = 2 pi int_{-1}^{1} C e^{kappa (t - 1)} dt
    # with t = 1 - u / kappa (u in [0, 2 kappa]) the integrand is C e^{-u} / kappa at any kappa;
    # beyond u = 60 it is below e^-60 of its peak
    u = np.linspace(0.0, min(2 * kappa, 60.0), 400001)
    t = 1 - u / kappa
    f = G.vmf_norm(kappa) * np.exp(kappa * (t - 1)) / kappa
    integral = 2 * np.pi * np.trapezoid(f, u)
    assert integral == pytest.approx(1.0, rel=1e-6)
    if kappa < 50:
        assert G.vmf_norm(kappa) == pytest.approx(kappa / (4 * np.pi * np.sinh(kappa)) * np.exp(kappa), rel=1e-12)


def test_mixture_integrates_to_one():
    rng = np.random.default_rng(1)
    raw = _rand_raw(rng, 3, logk=(-2.0, 2.5))
    # Gauss-Legendre in cos(theta) x uniform in phi
    ct, wt = np.polynomial.legendre.leggauss(400)
    ph = (np.arange(800) + 0.5) * 2 * np.pi / 800
    C, P = np.meshgrid(ct, ph, indexing="ij")
    st = np.sqrt(1 - C ** 2)
    d = np.stack([st * np.cos(P), st * np.sin(P), C], -1).reshape(-1, 3)
    w = (wt[:, None] * np.full_like(P, 2 * np.pi / 800)).ravel()
    for r in raw:
        q = G.mixture_pdf(np.repeat(r[None], len(d), 0), d)
        assert np.sum(q * w) == pytest.approx(1.0, rel=1e-6)


def _chi2_sphere(dirs, pdf_fn, nb_t=16, nb_p=16):
    """Chi-square of directions against a density on (cos theta, phi) bins."""
    ct = np.clip(dirs[:, 2], -1, 1)
    ph = np.mod(np.arctan2(dirs[:, 1], dirs[:, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(ct, ph, bins=[nb_t, nb_p], range=[[-1, 1], [0, 2 * np.pi]])
    # expected: Gauss-Legendre within each bin
    gx, gw = np.polynomial.legendre.leggauss(12)
    exp = np.zeros((nb_t, nb_p))
    for a in range(nb_t):
        c0, c1 = -1 + 2 * a / nb_t, -1 + 2 * (a + 1) / nb_t
        cs = 0.5 * (c1 - c0) * gx + 0.5 * (c1 + c0)
        for b in range(nb_p):
            p0, p1 = 2 * np.pi * b / nb_p, 2 * np.pi * (b + 1) / nb_p
            ps = 0.5 * (p1 - p0) * gx + 0.5 * (p1 + p0)
            C, P = np.meshgrid(cs, ps, indexing="ij")
            s = np.sqrt(1 - C ** 2)
            d = np.stack([s * np.cos(P), s * np.sin(P), C], -1).reshape(-1, 3)
            W = np.outer(gw, gw).ravel() * 0.25 * (c1 - c0) * (p1 - p0)
            exp[a, b] = np.sum(pdf_fn(d) * W)
    exp *= len(dirs)
    m = exp > 5
    chi = np.sum((obs[m] - exp[m]) ** 2 / exp[m])
    return stats.chi2.sf(chi, m.sum() - 1), obs[~m].sum(), exp[~m].sum()


def test_sampler_chi_square():
    rng = np.random.default_rng(2)
    raw = _rand_raw(rng, 1, logk=(0.0, 3.0))[0]
    n = 400000
    d = G.mixture_sample(np.repeat(raw[None], n, 0), rng.random((n, 3)))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
    p, _, _ = _chi2_sphere(d, lambda x: G.mixture_pdf(np.repeat(raw[None], len(x), 0), x))
    assert p > 1e-3


def test_sampler_chi_square_fails_for_a_planted_mistake():
    """A sampler with the vMF inverse CDF's sign flipped must fail the test."""
    rng = np.random.default_rng(3)
    raw = _rand_raw(rng, 1, logk=(1.0, 2.0))[0]
    n = 200000
    d = G.mixture_sample(np.repeat(raw[None], n, 0), rng.random((n, 3)))
    _, mu, _ = G.mixture(raw[None])
    # reflect each sample through the plane orthogonal to the lobe mean that generated it: approximate by the
    # first lobe, which changes the density unless every lobe is symmetric
    m = mu[0, 0]
    d_bad = d - 2 * (d @ m)[:, None] * m[None]
    p, _, _ = _chi2_sphere(d_bad, lambda x: G.mixture_pdf(np.repeat(raw[None], len(x), 0), x))
    assert p < 1e-6


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(4)
    raw = _rand_raw(rng, 6)
    wi = _unit(rng, 6)
    wt = rng.uniform(0.5, 2.0, 6)
    g = G.loss_grad(raw, wi, wt)
    h = 1e-6
    for f in range(40):
        rp, rm = raw.copy(), raw.copy()
        rp[:, f] += h
        rm[:, f] -= h
        fd = (G.loss(rp, wi, wt) - G.loss(rm, wi, wt)) / (2 * h)
        assert np.allclose(g[:, f], fd, rtol=1e-5, atol=1e-7), f
    assert np.all(g[:, 40:] == 0)


def test_loss_gradient_zero_outside_kappa_clamp():
    rng = np.random.default_rng(5)
    raw = _rand_raw(rng, 2)
    raw[0, 4] = 11.0   # above the clamp
    raw[1, 9] = -5.0   # below
    g = G.loss_grad(raw, _unit(rng, 2), np.ones(2))
    assert g[0, 4] == 0 and g[1, 9] == 0


@pytest.mark.parametrize("H", [64, 128])
def test_backward_finite_differences(H):
    rng = np.random.default_rng(6 + H)
    n = 12
    params = rng.normal(0, 0.15, G.num_params(H))
    Ws, bs = G.unpack_params(params, H)
    Ws[0][~G.input_used()] = 0.0
    params = G.pack_params(Ws, bs)
    lo, inv = G.guide_box([-1, -1, -1], [1, 1, 1])
    feats = G.encode(rng.uniform(-1, 1, (n, 3)).astype(np.float32), _unit(rng, n), _unit(rng, n), lo, inv)
    wi = _unit(rng, n)
    wt = rng.uniform(0.2, 1.0, n)
    L, g = G.backward(params, H, feats, wi, wt, bf16_emul=False)

    def total(p):
        raw, _ = G.forward(p, H, feats, bf16_emul=False)
        return np.sum(G.loss(raw, wi, wt))

    assert L == pytest.approx(total(params), rel=1e-12)
    Wsz = [fi * fo for fi, fo in G.layer_dims(H)]
    offs = np.cumsum([0] + Wsz)
    idx = []
    for l in range(5):   # a few weights of every layer, plus biases
        idx += list(rng.integers(offs[l], offs[l + 1], 6))
    idx += list(rng.integers(offs[5], params.size, 10))
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (total(pp) - total(pm)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-8), i


def test_adam_first_step_is_lr_times_sign():
    p, m, v = G.adam(np.array([1.0, -2.0]), np.array([0.3, -5.0]), np.zeros(2), np.zeros(2), 1, lr=1e-3)
    assert p == pytest.approx([1.0 - 1e-3, -2.0 + 1e-3], abs=1e-9)
    # second step with the same gradient: m-hat = g, v-hat = g^2 again
    p2, _, _ = G.adam(p, np.array([0.3, -5.0]), m, v, 2, lr=1e-3)
    assert p2 == pytest.approx([1.0 - 2e-3, -2.0 + 2e-3], abs=1e-8)
