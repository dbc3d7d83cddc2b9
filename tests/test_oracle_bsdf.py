"""Pins of the oracle's BSDF restatement (oracle/refcpu/ref_bsdf.hpp, after
src/material/disney.cpp, microfacet.cpp, reflection.hpp, diffuse_bxdf.cpp,
phase_function.hpp) by closed forms and by consistency between sample() and
pdf() — none of which retypes the formula under test:

* dielectric Fresnel against the textbook s/p-polarised Fresnel equations,
  0.04 at normal incidence for eta 1.5, 1 under total internal reflection;
* GGX / GTR1 normalisation: integral of D(h) cos(theta_h) over the hemisphere = 1;
* Heitz's visible-normal identity: integral of G1(wo) max(0, wo.h) D(h) / wo.z = 1;
* Lambert f = albedo / pi and integral of f cos = albedo;
* reciprocity of the Disney outside-reflection lobes;
* Henyey-Greenstein integral 1 and mean cosine g;
* chi-square of sample() histograms against integrated pdf() (fixed mode);
  the reference mode fails it (defect Q2), which shows the test has power.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from tests.materials import c2_materials
from tracer_amd.abi import TR_MAT_DISNEY, TR_MAT_LAMBERT, make_material


def _lib():
    from oracle import refcpu
    L = refcpu.lib()
    d = C.c_double
    for n, k in (("ref_fresnel", 3), ("ref_D_gtr2", 5), ("ref_D_gtr1", 2), ("ref_G1_gtr2", 5)):
        f = getattr(L, n)
        f.argtypes = [d] * k
        f.restype = d
    return L


def fresnel_textbook(n1, n2, ci):
    si = np.sqrt(max(0.0, 1 - ci * ci))
    st = n1 / n2 * si
    if st >= 1:
        return 1.0
    ct = np.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_known_answers():
    L = _lib()
    assert abs(L.ref_fresnel(1.5, 1.0, 1.0) - 0.04) < 1e-15
    # reference argument order: dielectric_fresnel(eta_i, eta_o, cos) with sin_t = eta_o/eta_i sin_i
    for ci in np.linspace(0.01, 1.0, 37):
        assert abs(L.ref_fresnel(1.5, 1.0, ci) - fresnel_textbook(1.0, 1.5, ci)) < 1e-12
        assert abs(L.ref_fresnel(1.0, 1.5, ci) - fresnel_textbook(1.5, 1.0, ci)) < 1e-12
    assert L.ref_fresnel(1.0, 1.5, 0.3) == 1.0     # total internal reflection
    assert abs(L.ref_fresnel(1.5, 1.0, -0.7) - L.ref_fresnel(1.0, 1.5, 0.7)) < 1e-15


def _hemisphere_grid(n_t=1500, n_p=256, power=3):
    # theta concentrated near the pole (sharp lobes), midpoint rule
    s = (np.arange(n_t) + 0.5) / n_t
    th = 0.5 * np.pi * s ** power
    dth = 0.5 * np.pi * power * s ** (power - 1) / n_t
    ph = (np.arange(n_p) + 0.5) / n_p * 2 * np.pi
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    dw = (np.sin(T) * dth[:, None]) * (2 * np.pi / n_p)
    h = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], -1)
    return h.reshape(-1, 3), dw.reshape(-1)


@pytest.mark.parametrize("ax,ay", [(0.05, 0.05), (0.2, 0.2), (0.5, 0.1), (1.0, 1.0)])
def test_gtr2_normalisation(ax, ay):
    L = _lib()
    h, dw = _hemisphere_grid()
    D = np.array([L.ref_D_gtr2(*v, ax, ay) for v in h])
    assert abs((D * h[:, 2] * dw).sum() - 1.0) < 2e-3


@pytest.mark.parametrize("alpha", [0.0001, 0.01, 0.1])
def test_gtr1_normalisation(alpha):
    L = _lib()
    h, dw = _hemisphere_grid(3000, 8, power=4)
    D = np.array([L.ref_D_gtr1(v[2], alpha) for v in h])
    assert abs((D * h[:, 2] * dw).sum() - 1.0) < 3e-3


@pytest.mark.parametrize("wo_theta", [0.3, 1.0, 1.4])
def test_visible_normal_identity(wo_theta):
    """Integral over h of G1(wo) max(0, wo.h) D(h) / wo.z = 1 (Heitz 2014)."""
    L = _lib()
    ax, ay = 0.3, 0.15
    wo = np.array([np.sin(wo_theta) * np.cos(0.4), np.sin(wo_theta) * np.sin(0.4), np.cos(wo_theta)])
    h, dw = _hemisphere_grid(1200, 256, power=2)
    D = np.array([L.ref_D_gtr2(*v, ax, ay) for v in h])
    g1 = L.ref_G1_gtr2(*wo, ax, ay)
    val = (g1 * np.maximum(0, h @ wo) * D / wo[2] * dw).sum()
    assert abs(val - 1.0) < 3e-3


def test_lambert_closed_form():
    from oracle import refcpu
    rng = np.random.default_rng(0)
    mats = [make_material(type=TR_MAT_LAMBERT, base_color=(0.3, 0.6, 0.9))]
    n = 4096
    wi = rng.normal(size=(n, 3)); wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    ids = np.zeros(n, np.int32)
    f, pdf = refcpu.bsdf_eval(mats, ids, wi.astype(np.float32), wo.astype(np.float32))
    up = (wi[:, 2] > 0) & (wo[:, 2] > 0)
    assert np.allclose(f[up], np.array([0.3, 0.6, 0.9]) / np.pi, rtol=1e-6)
    assert np.all(f[~up] == 0)
    assert np.allclose(pdf[up], wi[up, 2].astype(np.float32) / np.pi, rtol=1e-6)
    # integral of f cos over the hemisphere = albedo
    h, dw = _hemisphere_grid(400, 64, power=1)
    f2, _ = refcpu.bsdf_eval(mats, np.zeros(len(h), np.int32), h.astype(np.float32),
                             np.tile(np.float32([0, 0, 1]), (len(h), 1)))
    assert np.allclose((f2 * h[:, 2:3] * dw[:, None]).sum(0), [0.3, 0.6, 0.9], rtol=2e-4)


def test_disney_reflection_is_reciprocal():
    from oracle import refcpu
    rng = np.random.default_rng(1)
    mats = [m for m in c2_materials(2, 24) if m.type == TR_MAT_DISNEY]
    n = 1 << 13
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    a = rng.normal(size=(n, 3)); a[:, 2] = np.abs(a[:, 2]) + 0.05; a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.normal(size=(n, 3)); b[:, 2] = np.abs(b[:, 2]) + 0.05; b /= np.linalg.norm(b, axis=1, keepdims=True)
    a, b = a.astype(np.float32), b.astype(np.float32)
    fab, _ = refcpu.bsdf_eval(mats, ids, a, b)
    fba, _ = refcpu.bsdf_eval(mats, ids, b, a)
    assert np.allclose(fab, fba, rtol=1e-10, atol=1e-14)


def test_hg_normalisation_and_mean_cosine():
    from oracle import refcpu
    L = refcpu.lib()
    s = np.linspace(0, 1, 20001)
    uu = 1 - 2 * s ** 3          # dense near the forward peak at u = 1
    for g in (-0.5, 0.0, 0.3, 0.7, 0.95):
        p = np.array([L.ref_hg_eval(g, -x) for x in uu])
        assert abs(2 * np.pi * -np.trapezoid(p, uu) - 1) < 1e-4
        # the phase argument u = -cos(wi, wo) is the cosine of the scattering angle; its mean is g
        assert abs(2 * np.pi * -np.trapezoid(uu * p, uu) - g) < 1e-4


def _chi2_sample_vs_pdf(mat, wo, fixed, n=200_000, seed=0, nt=16, nphi=24):
    """Histogram of sample() over (cos theta, phi) cells against n * integral of pdf() per cell,
    plus one cell for invalid samples (expected n * (1 - integral of pdf))."""
    from oracle import refcpu
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wos = np.tile(np.float32(wo), (n, 1))
    wi, f, pdf = refcpu.bsdf_sample([mat], np.zeros(n, np.int32), wos, u, precision=1, fixed=fixed)
    ok = pdf > 0
    ct = np.clip(wi[ok, 2], -1, 1)
    ph = np.mod(np.arctan2(wi[ok, 1], wi[ok, 0]), 2 * np.pi)
    obs, _, _ = np.histogram2d(ct, ph, bins=[nt, nphi], range=[[-1, 1], [0, 2 * np.pi]])
    # pdf integrated per cell with an 8x8 midpoint sub-grid
    k = 8
    cg = -1 + (np.arange(nt * k) + 0.5) * (2 / (nt * k))
    pg = (np.arange(nphi * k) + 0.5) * (2 * np.pi / (nphi * k))
    CC, PP = np.meshgrid(cg, pg, indexing="ij")
    sg = np.sqrt(1 - CC ** 2)
    dirs = np.stack([sg * np.cos(PP), sg * np.sin(PP), CC], -1).reshape(-1, 3).astype(np.float32)
    _, pv = refcpu.bsdf_eval([mat], np.zeros(len(dirs), np.int32), dirs, np.tile(np.float32(wo), (len(dirs), 1)),
                             precision=1, fixed=fixed)
    cell = (2 / (nt * k)) * (2 * np.pi / (nphi * k))
    exp = (pv.reshape(nt * k, nphi * k) * cell).reshape(nt, k, nphi, k).sum(axis=(1, 3)) * n
    obs = np.append(obs.ravel(), (~ok).sum())
    exp = np.append(exp.ravel(), max(n - exp.sum(), 0.0))
    keep = exp > 20
    chi2 = ((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum()
    dof = keep.sum() - 1
    return stats.chi2.sf(chi2, dof), obs[~keep].sum() / n


# clearcoat_gloss 0 (GTR1 alpha 0.1): at gloss 1 (alpha 1e-4) the clearcoat lobe
# is a near-delta spike that the per-cell pdf quadrature below cannot resolve;
# that lobe is pinned by test_gtr1_normalisation instead.
CHI2_CASES = [
    dict(type=TR_MAT_DISNEY, base_color=(0.8, 0.5, 0.3), roughness=0.5, clearcoat=0.5, clearcoat_gloss=0.0, sheen=0.3),
    dict(type=TR_MAT_DISNEY, base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.35),
    dict(type=TR_MAT_DISNEY, base_color=(0.4, 0.7, 0.5), roughness=0.6, transmission=1.0, transmission_roughness=0.6),
    dict(type=TR_MAT_DISNEY, base_color=(0.2, 0.3, 0.8), roughness=0.4, anisotropic=0.7, clearcoat=1.0,
         clearcoat_gloss=0.0),
]


@pytest.mark.parametrize("case", range(len(CHI2_CASES)))
@pytest.mark.parametrize("wo", [(0.3, 0.2, 0.93), (0.8, -0.1, 0.59)])
def test_fixed_sampling_matches_pdf(case, wo):
    wo = np.array(wo) / np.linalg.norm(wo)
    p, lost = _chi2_sample_vs_pdf(make_material(**CHI2_CASES[case]), wo, fixed=1)
    assert p > 1e-4, p
    assert lost < 0.02


def test_fixed_sampling_matches_pdf_inside():
    """lwo.z < 0 (inside a transmissive object): macro-Fresnel split."""
    wo = np.array([0.2, 0.3, -0.93]) / np.linalg.norm([0.2, 0.3, -0.93])
    p, lost = _chi2_sample_vs_pdf(make_material(**CHI2_CASES[2]), wo, fixed=1)
    assert p > 1e-4, p


def test_reference_sampling_fails_chi_square():
    """Defect Q2: the reference's sample() never keeps the diffuse lobe it
    draws, so its sampling density differs from pdf()."""
    wo = np.array([0.3, 0.2, 0.93]) / np.linalg.norm([0.3, 0.2, 0.93])
    p, _ = _chi2_sample_vs_pdf(make_material(**CHI2_CASES[0]), wo, fixed=0)
    assert p < 1e-6


def test_float_restatement_close_to_fp64():
    """The reference-precision (float) restatement stays close to its FP64 twin
    away from grazing angles (a report of the reference's own FP32 error)."""
    from oracle import refcpu
    rng = np.random.default_rng(5)
    mats = c2_materials(6, 16)
    n = 1 << 14
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wi = rng.normal(size=(n, 3)).astype(np.float32); wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3)).astype(np.float32); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    f32, p32 = refcpu.bsdf_eval(mats, ids, wi, wo, precision=0, fixed=1)
    f64, p64 = refcpu.bsdf_eval(mats, ids, wi, wo, precision=1, fixed=1)
    dom = (np.abs(wi[:, 2]) > 0.05) & (np.abs(wo[:, 2]) > 0.05)
    m = dom[:, None] & (f64 > 1e-4)
    rel = np.abs(f32 - f64)[m] / f64[m]
    print(f"reference-precision Disney f: median rel err {np.median(rel):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 1e-5


# ---- Phong (src/material/phong.cpp, phong_specular_bxdf.cpp) -------------------
PHONG_CASES = [
    dict(base_color=(0.5, 0.3, 0.2), specular_scale=(0.6, 0.2, 0.1), shininess=1.0),    # kd + ks > 1: normalised by 1.1
    dict(base_color=(0.2, 0.4, 0.3), specular_scale=(0.3, 0.3, 0.3), shininess=12.0),
    dict(base_color=(0.05, 0.05, 0.05), specular_scale=(0.8, 0.7, 0.6), shininess=60.0),
]


def _phong(**kw):
    from tracer_amd.abi import TR_MAT_PHONG
    return make_material(type=TR_MAT_PHONG, **kw)


def test_phong_with_black_specular_is_lambert():
    """ks = 0 makes the aggregate a single diffuse lobe of weight 1: the Phong
    restatement then equals the Lambert one value for value and, in reference
    mode, sample for sample (fixed mode re-evaluates the sampled direction,
    so it agrees to rounding)."""
    from oracle import refcpu
    rng = np.random.default_rng(20)
    kd = (0.3, 0.6, 0.9)
    mats = [_phong(base_color=kd, specular_scale=0.0, shininess=7.0), make_material(type=TR_MAT_LAMBERT, base_color=kd)]
    n = 4096
    wi = rng.normal(size=(n, 3)).astype(np.float32)
    wo = rng.normal(size=(n, 3)).astype(np.float32)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    for fixed in (0, 1):
        a = refcpu.bsdf_eval(mats, np.zeros(n, np.int32), wi, wo, fixed=fixed)
        b = refcpu.bsdf_eval(mats, np.ones(n, np.int32), wi, wo, fixed=fixed)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        sa = refcpu.bsdf_sample(mats, np.zeros(n, np.int32), wo, u, fixed=fixed)
        sb = refcpu.bsdf_sample(mats, np.ones(n, np.int32), wo, u, fixed=fixed)
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y) if fixed == 0 else np.allclose(x, y, rtol=1e-12, atol=0)


@pytest.mark.parametrize("case", range(len(PHONG_CASES)))
def test_phong_albedo_at_normal_incidence(case):
    """For wo = n the specular lobe is centred on n, so the directional albedo
    has the closed form kd' + ks' (ns + 1) / (ns + 2), with kd', ks' divided by
    max(1, max_i(kd_i + ks_i)) (phong.cpp:36-41); the pdf integrates to
    w_d + w_s = 1 over the hemisphere."""
    from oracle import refcpu
    c = PHONG_CASES[case]
    kd, ks, ns = np.array(c["base_color"]), np.array(c["specular_scale"]), c["shininess"]
    dem = max(1.0, (kd + ks).max())
    want = kd / dem + ks / dem * (ns + 1) / (ns + 2)
    h, dw = _hemisphere_grid(1500, 32, power=3)
    wo = np.tile(np.float32([0, 0, 1]), (len(h), 1))
    f, pdf = refcpu.bsdf_eval([_phong(**c)], np.zeros(len(h), np.int32), h.astype(np.float32), wo)
    assert np.allclose((f * h[:, 2:3] * dw[:, None]).sum(0), want, rtol=1e-3)
    assert abs((pdf * dw).sum() - 1.0) < 1e-3


def test_phong_lobe_weights_follow_luminance():
    """Away from the specular lobe the pdf is w_d cos / pi with
    w_d = lum(kd) / (lum(kd) + lum(ks)), lum = 0.2126 R + 0.7152 G + 0.0722 B
    (phong.cpp:43-53, spectrum.hpp:94-96); both black gives 1/2."""
    from oracle import refcpu
    wo = np.float32([[0.0, 0.0, 1.0]])
    wi = np.float32([[0.96, 0.0, 0.28]])   # 73.7 degrees from the lobe axis: cos^200 underflows to ~1e-112
    for kd, ks, wd in (((0.5, 0.3, 0.2), (0.1, 0.6, 0.3), None), ((0, 0, 0), (0, 0, 0), 0.5)):
        if wd is None:
            lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
            wd = lum(kd) / (lum(kd) + lum(ks))
        _, pdf = refcpu.bsdf_eval([_phong(base_color=kd, specular_scale=ks, shininess=200.0)], np.zeros(1, np.int32),
                                  wi, wo)
        assert abs(pdf[0] / (0.28 / np.pi) - wd) < 1e-6   # float32 inputs


@pytest.mark.parametrize("case", range(len(PHONG_CASES)))
@pytest.mark.parametrize("wo", [(0.3, 0.2, 0.93), (0.8, -0.1, 0.59)])
def test_phong_fixed_sampling_matches_pdf(case, wo):
    wo = np.array(wo) / np.linalg.norm(wo)
    p, lost = _chi2_sample_vs_pdf(_phong(**PHONG_CASES[case]), wo, fixed=1)
    assert p > 1e-4, p


# ---- Glass (src/material/glass.cpp) -----------------------------------------------
def _glass(kr=(1.0, 1.0, 1.0), kt=(1.0, 1.0, 1.0), ior=1.5):
    from tracer_amd.abi import TR_MAT_GLASS
    return make_material(type=TR_MAT_GLASS, base_color=kr, specular_scale=kt, ior=ior)


@pytest.mark.parametrize("theta_deg,inside", [(0.0, False), (35.0, False), (70.0, False), (20.0, True), (50.0, True)])
def test_glass_sampling_fresnel_snell(theta_deg, inside):
    """The reflected fraction is the textbook unpolarised Fresnel reflectance,
    reflected directions are mirror images, refracted ones obey
    n_i sin(theta_i) = n_t sin(theta_t) in the plane of incidence, and beyond
    the critical angle (from inside, 41.8 degrees for n = 1.5) every sample
    reflects.  eval and pdf vanish for every pair of directions."""
    from oracle import refcpu
    th = np.radians(theta_deg)
    wo = np.array([np.sin(th) * np.cos(0.3), np.sin(th) * np.sin(0.3), np.cos(th) * (-1 if inside else 1)])
    n = 40000
    rng = np.random.default_rng(41)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wos = np.tile(np.float32(wo), (n, 1))
    mat = _glass(kr=(0.9, 0.8, 0.7), kt=(0.64, 0.81, 1.0))
    wi, f, pdf = refcpu.bsdf_sample([mat], np.zeros(n, np.int32), wos, u, precision=1, fixed=1)
    n1, n2 = (1.5, 1.0) if inside else (1.0, 1.5)
    F = fresnel_textbook(n1, n2, np.cos(th))
    refl = wi[:, 2] * wo[2] > 0
    k = refl.sum()
    assert abs(k / n - F) < 5 * np.sqrt(F * (1 - F) / n) + 1e-12
    mirror = np.array([-wo[0], -wo[1], wo[2]])
    assert np.allclose(wi[refl], mirror, atol=1e-6)
    assert np.allclose(pdf[refl], F, rtol=1e-6)
    # reflected weight f |cos| / pdf = kr
    assert np.allclose(f[refl] * np.abs(wi[refl, 2:3]) / pdf[refl, None], [0.9, 0.8, 0.7], rtol=1e-6)
    if (~refl).any():
        t = wi[~refl]
        sin_i = np.sin(th)
        sin_t = np.linalg.norm(t[:, :2], axis=1)
        assert np.allclose(n1 * sin_i, n2 * sin_t, atol=1e-6)
        if theta_deg > 0:   # in the plane of incidence, on the other side of the normal
            az = np.arctan2(t[:, 1], t[:, 0])
            assert np.abs(np.angle(np.exp(1j * (az - 0.3 - np.pi)))).max() < 1e-5
        # refracted weight: (n_i / n_t)^2 sqrt(kt)
        wgt = f[~refl] * np.abs(t[:, 2:3]) / pdf[~refl, None]
        assert np.allclose(wgt, (n1 / n2) ** 2 * np.sqrt([0.64, 0.81, 1.0]), rtol=1e-6)
    crit = np.degrees(np.arcsin(1 / 1.5))
    if inside and theta_deg > crit:
        assert k == n
    fe, pe = refcpu.bsdf_eval([mat], np.zeros(n, np.int32), wi.astype(np.float32), wos)
    assert np.all(fe == 0) and np.all(pe == 0)
