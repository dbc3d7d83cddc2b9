"""Pins of the oracle's image environment light (oracle/refcpu/ref_env.hpp,
after src/light/environment_light.cpp:20-196 and
src/utility/distribution.hpp:12-110):

* a constant image integrates to L * 4 pi (the texel solid angles tile the
  sphere), and its samples have E[1 / pdf] = 4 pi;
* the pdf integrates to 1 over the sphere, and the histogram of sampled
  directions matches it cell by cell (chi-square); the reference's pdf()
  (defect Q28, reference mode) integrates to the marginal integral squared,
  which the same check rejects;
* a rotation of the environment rotates the radiance it returns.
"""
import numpy as np
import pytest
from scipy import stats


def _env(rgb, rot=None):
    from oracle import refcpu
    return refcpu.RefEnv(rgb, rot)


def _sphere_grid(nt=400, nphi=800):
    t = (np.arange(nt) + 0.5) * np.pi / nt
    p = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / nt) * (2 * np.pi / nphi)).reshape(-1)
    return d, dw


def _blob_image(w=48, h=24, seed=3):
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    U, V = np.meshgrid(u, v)
    img = 0.2 + 0.3 * rng.uniform(size=(h, w, 3))
    img += 40.0 * np.exp(-((U - 0.3) ** 2 + (V - 0.35) ** 2) / 0.002)[..., None] * np.array([1.0, 0.8, 0.6])
    img[h - 3:] = 0.0   # a black band near the bottom pole
    return img.astype(np.float32)


def test_constant_image_integral_and_pdf():
    L = 0.7
    e = _env(np.full((8, 16, 3), L, np.float32))
    assert np.allclose(e.radiance_integral(), L * 4 * np.pi, rtol=1e-5)
    rng = np.random.default_rng(0)
    wi, pdf, rad = e.sample(rng.uniform(size=(20000, 2)))
    ok = pdf > 0
    assert ok.mean() > 0.999
    assert np.allclose(rad[ok], L, rtol=1e-6)
    # E[1 / pdf] is the measure of the support: 4 pi for an image black nowhere
    assert abs(np.mean(1 / (4 * np.pi * pdf[ok])) - 1) < 0.01
    # pdf() at the sampled direction agrees away from the poles, where acos of
    # a float cosine near 1 loses the angle's digits
    _, pe = e.eval(wi[ok])
    away = np.abs(wi[ok, 2]) < 0.999
    assert np.mean(np.isclose(pe[away], pdf[ok][away], rtol=1e-4)) > 0.999


@pytest.mark.parametrize("quirks,expect_one", [(False, True), (True, False)])
def test_pdf_integrates_to_one(quirks, expect_one):
    e = _env(_blob_image())
    d, dw = _sphere_grid()
    _, pdf = e.eval(d.astype(np.float32), quirks=quirks)
    total = float((pdf * dw).sum())
    print(f"quirks={quirks}: integral of pdf {total:.5f}")
    if expect_one:
        assert abs(total - 1) < 5e-3
    else:
        assert abs(total - 1) > 0.05


def test_sampling_matches_pdf_chi_square():
    e = _env(_blob_image())
    n = 400_000
    rng = np.random.default_rng(1)
    wi, pdf, _ = e.sample(rng.uniform(size=(n, 2)))
    assert np.all(pdf > 0)
    nt, nphi = 24, 48
    th = np.arccos(np.clip(wi[:, 2], -1, 1))
    ph = np.mod(np.arctan2(wi[:, 1], wi[:, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(th, ph, bins=[nt, nphi], range=[[0, np.pi], [0, 2 * np.pi]])
    # expected counts: the pdf integrated over each (theta, phi) cell on a fine sub-grid
    k = 8
    t = (np.arange(nt * k) + 0.5) * np.pi / (nt * k)
    p = (np.arange(nphi * k) + 0.5) * 2 * np.pi / (nphi * k)
    T, P = np.meshgrid(t, p, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    _, pe = e.eval(d)
    cell = (np.sin(T) * (np.pi / (nt * k)) * (2 * np.pi / (nphi * k))).reshape(-1)
    exp = (pe * cell).reshape(nt, k, nphi, k).sum(axis=(1, 3)) * n
    keep = exp > 20
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    pval = stats.chi2.sf(chi2, keep.sum() - 1)
    print(f"chi2 {chi2:.1f} dof {keep.sum() - 1} p {pval:.3g}; mass in kept cells {exp[keep].sum() / n:.4f}")
    assert pval > 1e-4
    assert obs[~keep].sum() < 0.01 * n


def test_rotation_rotates_radiance():
    img = _blob_image()
    a = np.radians(35.0)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    e0, e1 = _env(img), _env(img, R)
    rng = np.random.default_rng(2)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    L1, p1 = e1.eval(d.astype(np.float32))
    L0, p0 = e0.eval((d @ R.T).astype(np.float32))
    assert np.allclose(L1, L0, rtol=2e-4, atol=1e-5)
    assert np.allclose(p1, p0, rtol=2e-4, atol=1e-9)
