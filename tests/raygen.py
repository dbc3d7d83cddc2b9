"""Seeded ray batches shared by the traversal tests (numpy only)."""
from __future__ import annotations

import numpy as np

BIG = np.float32(1e30)


def scene_bounds(scene):
    P = scene.positions.reshape(-1, 3)
    return P.min(axis=0), P.max(axis=0)


def unit_sphere(rng, n):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def random_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = scene_bounds(scene)
    o = rng.uniform(lo, hi, (n, 3))
    d = unit_sphere(rng, n)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0
    r[:, 4:7] = d
    r[:, 7] = BIG
    return r


def interior_rays(scene, n, seed, margin=1e-3):
    """Rays aimed at interior points of random triangles (barycentrics >= margin)."""
    rng = np.random.default_rng(seed)
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    I = scene.indices.reshape(-1, 3)
    t = rng.integers(0, I.shape[0], n)
    w = rng.dirichlet((1, 1, 1), n)
    w = margin + (1 - 3 * margin) * w
    target = (P[I[t, 0]] * w[:, :1] + P[I[t, 1]] * w[:, 1:2] + P[I[t, 2]] * w[:, 2:3])
    lo, hi = scene_bounds(scene)
    o = rng.uniform(lo, hi, (n, 3))
    d = target - o
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = BIG
    return r, t


def edge_case_rays(scene, n, seed):
    """Rays where a conservative box test matters: aimed exactly at triangle
    vertices and edge midpoints (shared by neighbouring triangles), axis-aligned
    directions (zero direction components), origins 1000 scene units away and
    origins on a vertex.  n rays, in five equal groups."""
    rng = np.random.default_rng(seed)
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    I = scene.indices.reshape(-1, 3)
    lo, hi = scene_bounds(scene)
    k = n // 5
    o = rng.uniform(lo, hi, (n, 3))
    tri = rng.integers(0, I.shape[0], n)
    corner = rng.integers(0, 3, n)
    vert = P[I[tri, corner]]
    mid = 0.5 * (P[I[tri, corner]] + P[I[tri, (corner + 1) % 3]])
    w = rng.dirichlet((1, 1, 1), n)
    interior = P[I[tri, 0]] * w[:, :1] + P[I[tri, 1]] * w[:, 1:2] + P[I[tri, 2]] * w[:, 2:3]
    d = np.empty((n, 3))
    d[:k] = vert[:k] - o[:k]
    d[k:2 * k] = mid[k:2 * k] - o[k:2 * k]
    a0 = rng.integers(0, 3, k)
    a1 = (a0 + rng.integers(1, 3, k)) % 3   # a different axis
    axes = np.eye(3)[a0] * rng.choice([-1.0, 1.0], (k, 1))
    two = rng.random(k) < 0.5   # half of them with two non-zero components
    axes = np.where(two[:, None], axes + np.eye(3)[a1] * rng.choice([-1.0, 1.0], (k, 1)), axes)
    d[2 * k:3 * k] = axes
    far = rng.normal(size=(k, 3))
    far /= np.linalg.norm(far, axis=1, keepdims=True)
    o[3 * k:4 * k] = 0.5 * (lo + hi) + 1000.0 * far
    d[3 * k:4 * k] = interior[3 * k:4 * k] - o[3 * k:4 * k]
    o[4 * k:] = vert[4 * k:]
    d[4 * k:] = rng.normal(size=(n - 4 * k, 3))
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = BIG
    return r


def hit_prims(h):
    return h[:, 1].copy().view(np.int32)


def exact_crossing(scene, rays, prims):
    """FP64 crossing of each ray with the plane of the given triangle:
    (min barycentric, |cos(ray, n_g)|, t, |o - A|, shortest edge)."""
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    I = scene.indices.reshape(-1, 3)
    v0, v1, v2 = P[I[prims, 0]], P[I[prims, 1]], P[I[prims, 2]]
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = (e1 * p).sum(1)
    tv = o - v0
    q = np.cross(tv, e1)
    u = (tv * p).sum(1) / det
    w = (d * q).sum(1) / det
    t = (e2 * q).sum(1) / det
    n = np.cross(e1, e2)
    cos = np.abs((n * d).sum(1)) / (np.linalg.norm(n, axis=1) * np.linalg.norm(d, axis=1))
    emin = np.minimum(np.minimum(np.linalg.norm(e1, axis=1), np.linalg.norm(e2, axis=1)),
                      np.linalg.norm(v2 - v1, axis=1))
    return np.minimum(np.minimum(u, w), 1 - u - w), cos, t * np.linalg.norm(d, axis=1), np.linalg.norm(tv, axis=1), emin


EPS32 = 2.0 ** -24


def well_conditioned(scene, rays, prims):
    """A hit is well conditioned when the ray crosses the triangle inside its
    edges by more than FP32 can resolve from the ray origin (exact barycentric
    margin >= max(1e-3, 16 eps |o - A| / shortest edge)), at |cos| >= 0.02,
    with t condition number |o - A| / (t cos) <= 50."""
    out = np.zeros(len(prims), bool)
    k = prims >= 0
    if k.any():
        mg, cs, tt, dist, emin = exact_crossing(scene, rays[k], prims[k])
        need = np.maximum(1e-3, 16 * EPS32 * dist / emin)
        out[k] = (mg >= need) & (cs >= 0.02) & (dist / np.maximum(tt * cs, 1e-30) <= 50)
    return out


def compare_with_mt(scene, rays, wt_hits, mt_hits):
    """Agreement of watertight closest hits with the reference's
    Moller-Trumbore closest hits (src/shape/triangle.cpp:145-175).

    On the well-conditioned suite (tests.raygen.well_conditioned holds for the
    watertight hit) both tests must name the same triangle unless the other
    test's triangle is itself not well conditioned (an edge or grazing hit
    that FP32 cannot resolve), and t must agree within 1e-5 relative.
    Everything else is reported, not asserted.  Returns a dict."""
    a, b = hit_prims(wt_hits), hit_prims(mt_hits)
    well = well_conditioned(scene, rays, a)
    other_well = well_conditioned(scene, rays, b)
    same = a == b
    genuine = well & ~same & (other_well | (b < 0))
    ok = well & same
    rel = np.abs(wt_hits[ok, 0] - mt_hits[ok, 0]) / np.abs(mt_hits[ok, 0])
    return {
        "well_fraction": float(well.mean()),
        "well_id_mismatch": int(genuine.sum()),
        "well_id_mismatch_fp32_unresolvable": int((well & ~same).sum() - genuine.sum()),
        "well_t_rel_max": float(rel.max()) if rel.size else 0.0,
        "all_id_mismatch_rate": float((~same).mean()),
    }
