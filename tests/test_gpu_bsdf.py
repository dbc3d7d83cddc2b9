"""GPU Disney / Lambert BSDF (tr_bsdf_eval_batch, tr_bsdf_sample_batch)
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
