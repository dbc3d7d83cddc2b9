"""GPU Disney / Lambert / Phong BSDF (tr_bsdf_eval_batch, tr_bsdf_sample_batch)
against the FP64 twin of the reference formulas (src/material/disney.cpp,
oracle/refcpu/ref_bsdf.hpp, defect corrections Q2/Q3/Q21/Q22 on).

Tolerance: 1e-5 relative on f (per channel) and pdf where f, pdf >= 1e-6,
on the well-conditioned domain of tests.materials.conditioned (every cosine
in the formulas >= 2e-2 and the Fresnel terms >= 2e-2 away from the critical
angle in cos^2; ~90% of uniformly random direction pairs).  Errors outside
the domain are printed, not asserted."""
import numpy as np
import pytest

from tests.materials import c2_materials, conditioned, sphere_dirs, tilted_normals

pytestmark = pytest.mark.gpu
REL_TOL = 1e-5


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def test_eval_pdf_matches_fp64(gpu_ctx):
    from oracle import refcpu
    rng = np.random.default_rng(10)
    mats = c2_materials(11, 48)
    n = 1 << 16
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wi, wo = sphere_dirs(rng, n), sphere_dirs(rng, n)
    f, pdf = gpu_ctx.bsdf_eval(mats, ids, wi, wo)
    fr, pr = refcpu.bsdf_eval(mats, ids, wi, wo, precision=1, fixed=1)
    dom = conditioned(wi, wo, mats, ids)
    fm = dom[:, None] & (fr >= 1e-6)
    pm = dom & (pr >= 1e-6)
    ef = _rel(f.astype(np.float64), fr)[fm]
    ep = _rel(pdf.astype(np.float64), pr)[pm]
    out_f = _rel(f.astype(np.float64), fr)[~dom[:, None] & (fr >= 1e-6)]
    print(f"domain {dom.mean():.3f}; f: n={ef.size} max={ef.max():.3g} p99.9={np.quantile(ef, 0.999):.3g}; "
          f"pdf: n={ep.size} max={ep.max():.3g}; outside domain f p99={np.quantile(out_f, 0.99):.3g}")
    # zero where the oracle is zero
    assert np.all(f[(fr == 0)] == 0)
    assert ef.max() <= REL_TOL
    assert ep.max() <= REL_TOL


def test_eval_with_tilted_geometric_normal(gpu_ctx):
    from oracle import refcpu
    rng = np.random.default_rng(12)
    mats = c2_materials(13, 16)
    n = 1 << 15
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wi, wo = sphere_dirs(rng, n), sphere_dirs(rng, n)
    ng = tilted_normals(rng, n)
    f, pdf = gpu_ctx.bsdf_eval(mats, ids, wi, wo, ng)
    fr, pr = refcpu.bsdf_eval(mats, ids, wi, wo, ng, precision=1, fixed=1)
    dom = conditioned(wi, wo, mats, ids) & (np.abs((wi * ng).sum(1)) >= 2e-2) & (np.abs((wo * ng).sum(1)) >= 2e-2)
    fm = dom[:, None] & (fr >= 1e-6)
    ef = _rel(f.astype(np.float64), fr)[fm]
    print(f"tilted f max rel {ef.max():.3g}")
    assert ef.max() <= REL_TOL
    pm = dom & (np.abs(pr) >= 1e-6)
    assert _rel(pdf.astype(np.float64), pr)[pm].max() <= REL_TOL


def test_sample_directions_match_fp64(gpu_ctx):
    """Same uniforms -> same sampled direction (non-fringe cases)."""
    from oracle import refcpu
    rng = np.random.default_rng(14)
    mats = c2_materials(15, 32)
    n = 1 << 15
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wo = sphere_dirs(rng, n)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wi, f, pdf = gpu_ctx.bsdf_sample(mats, ids, wo, u)
    wir, fr, pr = refcpu.bsdf_sample(mats, ids, wo, u, precision=1, fixed=1)
    both = (pdf > 0) & (pr > 0) & (np.abs(wo[:, 2]) >= 2e-2)
    agree = (pdf > 0) == (pr > 0)
    print(f"valid agreement {agree.mean():.6f}, both valid {both.sum()}")
    assert agree.mean() > 0.999
    ang = np.linalg.norm(wi[both] - wir[both], axis=1)
    print(f"direction abs diff: median {np.median(ang):.3g}, p99 {np.quantile(ang, 0.99):.3g}, max {ang.max():.3g}")
    assert np.quantile(ang, 0.99) < 1e-5
    assert np.quantile(ang, 0.9999) < 1e-3


def _phong_materials():
    from tracer_amd.abi import TR_MAT_PHONG, make_material
    rng = np.random.default_rng(31)
    mats = []
    for ns in (0.0, 1.0, 4.0, 12.0, 30.0, 80.0):
        for _ in range(3):
            mats.append(make_material(type=TR_MAT_PHONG, base_color=tuple(rng.uniform(0, 0.9, 3)),
                                      specular_scale=tuple(rng.uniform(0, 0.9, 3)), shininess=ns))
    mats.append(make_material(type=TR_MAT_PHONG, base_color=(0.6, 0.5, 0.4), specular_scale=0.0, shininess=10.0))
    mats.append(make_material(type=TR_MAT_PHONG, base_color=0.0, specular_scale=0.0, shininess=5.0))
    return mats


def test_phong_eval_pdf_matches_fp64(gpu_ctx):
    """Phong (phong.cpp, phong_specular_bxdf.cpp) against the FP64 oracle.
    Tolerance: 1e-5 relative plus the FP32 error of the lobe cosine c
    (<= 4e-7 absolute) carried through c^ns, i.e. ns * 4e-7 / c relative, on
    the domain where wi.z, wo.z >= 2e-2."""
    from oracle import refcpu
    rng = np.random.default_rng(32)
    mats = _phong_materials()
    n = 1 << 16
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wi, wo = sphere_dirs(rng, n), sphere_dirs(rng, n)
    f, pdf = gpu_ctx.bsdf_eval(mats, ids, wi, wo)
    fr, pr = refcpu.bsdf_eval(mats, ids, wi, wo, precision=1, fixed=1)
    wi64, wo64 = wi.astype(np.float64), wo.astype(np.float64)
    c = wi64[:, 2] * wo64[:, 2] - wi64[:, 0] * wo64[:, 0] - wi64[:, 1] * wo64[:, 1]
    ns = np.array([mats[i].shininess for i in ids])
    tol = REL_TOL + np.where(c > 0, ns * 4e-7 / np.maximum(c, 1e-30), 0.0)
    dom = (np.abs(wi[:, 2]) >= 2e-2) & (np.abs(wo[:, 2]) >= 2e-2)
    assert np.all(f[fr == 0] == 0) and np.all(pdf[pr == 0] == 0)
    fm = dom[:, None] & (fr >= 1e-6)
    pm = dom & (pr >= 1e-6)
    ef = (_rel(f.astype(np.float64), fr) / tol[:, None])[fm]
    ep = (_rel(pdf.astype(np.float64), pr) / tol)[pm]
    print(f"Phong: f n={ef.size} max err/tol {ef.max():.3g}; pdf n={ep.size} max err/tol {ep.max():.3g}")
    assert ef.max() <= 1.0 and ep.max() <= 1.0


def test_phong_sample_directions_match_fp64(gpu_ctx):
    """Same uniforms -> same lobe and direction (the lobe is built in the
    frame of Coord::create_from_z on both sides)."""
    from oracle import refcpu
    rng = np.random.default_rng(33)
    mats = _phong_materials()
    n = 1 << 15
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wo = sphere_dirs(rng, n)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wi, f, pdf = gpu_ctx.bsdf_sample(mats, ids, wo, u)
    wir, fr, pr = refcpu.bsdf_sample(mats, ids, wo, u, precision=1, fixed=1)
    agree = (pdf > 0) == (pr > 0)
    both = (pdf > 0) & (pr > 0) & (np.abs(wo[:, 2]) >= 2e-2)
    print(f"Phong sample: valid agreement {agree.mean():.6f}, both valid {both.sum()}, valid fraction {(pr > 0).mean():.3f}")
    assert agree.mean() > 0.999
    d = np.linalg.norm(wi[both] - wir[both], axis=1)
    print(f"direction abs diff: median {np.median(d):.3g}, p99 {np.quantile(d, 0.99):.3g}, max {d.max():.3g}")
    assert np.quantile(d, 0.99) < 1e-5
    assert np.quantile(d, 0.9999) < 1e-3


def test_glass_samples_match_fp64(gpu_ctx):
    """Glass (glass.cpp): the same uniform picks the same lobe, direction, f
    and pdf as the FP64 oracle.  Tolerance 2e-5 relative on f and pdf where
    |wo.z| >= 2e-2 and cos^2 of the refracted angle >= 2e-2: FP32 Fresnel
    from the same cosine is within 5.5e-6 there, and the GPU's FP32
    normalisation of wo moves the cosine by an ulp, which the Fresnel slope
    near the critical angle turns into ~1e-5."""
    from oracle import refcpu
    from tracer_amd.abi import TR_MAT_GLASS, make_material
    rng = np.random.default_rng(51)
    mats = [make_material(type=TR_MAT_GLASS, base_color=(0.95, 0.9, 0.8), specular_scale=(0.95, 0.9, 0.8), ior=1.5),
            make_material(type=TR_MAT_GLASS, base_color=(1.0, 1.0, 1.0), specular_scale=(0.5, 0.7, 0.9), ior=1.33),
            make_material(type=TR_MAT_GLASS, base_color=(0.7, 0.7, 0.7), specular_scale=(0.7, 0.7, 0.7), ior=2.4)]
    n = 1 << 15
    ids = rng.integers(0, len(mats), n).astype(np.int32)
    wo = sphere_dirs(rng, n)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    wi, f, pdf = gpu_ctx.bsdf_sample(mats, ids, wo, u)
    wir, fr, pr = refcpu.bsdf_sample(mats, ids, wo, u, precision=1, fixed=1)
    ior = np.array([mats[i].ior for i in ids])
    ci = np.abs(wo[:, 2].astype(np.float64))
    eta = np.where(wo[:, 2] > 0, 1 / ior, ior)
    ct2 = 1 - eta ** 2 * (1 - ci ** 2)
    dom = (ci >= 2e-2) & (np.abs(ct2) >= 2e-2)
    same = (pdf > 0) == (pr > 0)
    lobe = np.sign(wi[:, 2] * wo[:, 2]) == np.sign(wir[:, 2] * wo[:, 2])
    print(f"glass: valid agreement {same.mean():.6f}, lobe agreement {lobe[dom].mean():.6f}")
    assert same.mean() > 0.999 and lobe[dom].mean() > 0.999
    m = dom & lobe & (pr > 0)
    assert np.quantile(np.linalg.norm(wi[m] - wir[m], axis=1), 0.999) < 1e-5
    tol = 2e-5
    ep = _rel(pdf[m].astype(np.float64), pr[m])
    fm = m[:, None] & (fr > 0)
    ef = _rel(f.astype(np.float64), fr)[fm]
    print(f"glass: pdf max rel {ep.max():.3g}, f max rel {ef.max():.3g}")
    assert ep.max() <= tol and ef.max() <= tol
