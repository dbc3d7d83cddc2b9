"""Seeded material sets and direction batches for the BSDF tests."""
from __future__ import annotations

import numpy as np

from tracer_amd.abi import TR_MAT_DISNEY, TR_MAT_LAMBERT, make_material


def c2_materials(seed: int, n: int):
    """Disney parameters drawn like configuration 2, plus corner cases and a Lambert."""
    rng = np.random.default_rng(seed)
    mats = []
    for _ in range(n):
        r = float(rng.uniform(0.05, 1.0))
        mats.append(make_material(type=TR_MAT_DISNEY, base_color=tuple(rng.uniform(0, 1, 3)),
                                  metallic=float(rng.uniform() < 0.3), roughness=r, transmission=float(rng.uniform() < 0.1),
                                  transmission_roughness=r, ior=1.5, specular_scale=1.0, clearcoat=float(rng.uniform()),
                                  sheen=float(rng.uniform(0, 0.5)), sheen_tint=0.5, clearcoat_gloss=1.0))
    mats.append(make_material(type=TR_MAT_DISNEY, base_color=(0.8, 0.3, 0.2), roughness=0.05, transmission=1.0,
                              transmission_roughness=0.05, clearcoat=1.0, sheen=0.5))
    mats.append(make_material(type=TR_MAT_DISNEY, base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.3, anisotropic=0.6))
    mats.append(make_material(type=TR_MAT_DISNEY, base_color=(0.2, 0.6, 0.3), roughness=1.0, transmission=0.5,
                              transmission_roughness=0.7, specular_tint=0.7, sheen=0.3, sheen_tint=0.2, clearcoat=0.4,
                              clearcoat_gloss=0.3, ior=2.0))
    mats.append(make_material(type=TR_MAT_LAMBERT, base_color=(0.7, 0.5, 0.3)))
    return mats


def sphere_dirs(rng, n):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)


def tilted_normals(rng, n, max_tilt=0.3):
    t = rng.uniform(-max_tilt, max_tilt, (n, 2))
    g = np.concatenate([t, np.ones((n, 1))], axis=1)
    return (g / np.linalg.norm(g, axis=1, keepdims=True)).astype(np.float32)


def conditioned(wi, wo, mats, ids, th=2e-2):
    """Well-conditioned domain of the FP32 BSDF comparison: every cosine that
    enters the formulas (wi.n, wo.n, wi.h, wo.h, h.n for the half vector of
    the lobe in use) is at least th, and the Fresnel terms stay away from the
    critical angle (cos^2 of the transmitted angle >= th at the microfacet and,
    for wo inside, at the macro normal).  Outside it FP32 results from FP32
    inputs lose digits in proportion to 1/cos or 1/cos^2 for any evaluation
    order."""
    wi = wi.astype(np.float64)
    wo = wo.astype(np.float64)
    ior = np.array([max(1.01, mats[i].ior) for i in ids])
    same = wi[:, 2] * wo[:, 2] > 0
    hr = wi + wo
    hr /= np.linalg.norm(hr, axis=1, keepdims=True)
    hr *= np.where(wo[:, 2:3] < 0, -1.0, 1.0)
    eta = np.where(wo[:, 2] > 0, ior, 1 / ior)[:, None]
    ht = wo + wi * eta
    ht /= np.linalg.norm(ht, axis=1, keepdims=True)
    ht *= np.where(ht[:, 2:3] < 0, -1.0, 1.0)
    h = np.where(same[:, None], hr, ht)
    ci, cii = (wo * h).sum(1), (wi * h).sum(1)
    r = np.where(ci >= 0, 1 / ior, ior)
    ct2 = 1 - r * r * (1 - ci * ci)
    macro = np.where(wo[:, 2] < 0, 1 - ior * ior * (1 - wo[:, 2] ** 2), 1.0)
    return ((np.abs(wi[:, 2]) >= th) & (np.abs(wo[:, 2]) >= th) & (np.abs(ci) >= th) & (np.abs(cii) >= th)
            & (np.abs(h[:, 2]) >= th) & (np.abs(ct2) >= th) & (np.abs(macro) >= th))
