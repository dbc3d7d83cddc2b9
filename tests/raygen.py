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


def hit_prims(h):
    return h[:, 1].copy().view(np.int32)


def compare_with_mt(scene, rays, wt_hits, mt_hits, min_cos=1e-2):
    """Agreement of watertight hits with the reference Moller-Trumbore hits.

    Returns (id mismatch rate, fraction of mismatches that are edge cases,
    max relative t difference over hits with |cos(ray, n_g)| >= min_cos).  A
    mismatch is an edge case when the two triangles share a vertex, when
    either hit lies within 1e-3 (barycentric) of its triangle's edge — where
    Moller-Trumbore can leak through to a farther triangle — when either hit
    is grazing (|cos| < min_cos), or when the two t agree to 1e-5.  Below min_cos the FP32 t of any
    ray/triangle test is ill-conditioned (error ~ eps / cos)."""
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    I = scene.indices.reshape(-1, 3)
    a, b = hit_prims(wt_hits), hit_prims(mt_hits)
    mism = a != b
    adj = 0
    def near_edge(i, prim):
        # exact (FP64) barycentrics of the ray's crossing with the triangle's plane
        v = P[I[prim]]
        o = rays[i, 0:3].astype(np.float64)
        d = rays[i, 4:7].astype(np.float64)
        e1, e2 = v[1] - v[0], v[2] - v[0]
        p = np.cross(d, e2)
        det = e1 @ p
        if det == 0:
            return True
        tv = o - v[0]
        u = (tv @ p) / det
        w = (d @ np.cross(tv, e1)) / det
        return min(u, w, 1.0 - u - w) < 1e-3

    def grazing(i, prim):
        v = P[I[prim]]
        n = np.cross(v[1] - v[0], v[2] - v[0])
        d = rays[i, 4:7].astype(np.float64)
        return abs(n @ d) / (np.linalg.norm(n) * np.linalg.norm(d)) < min_cos

    for i in np.nonzero(mism)[0]:
        if (a[i] >= 0 and (near_edge(i, a[i]) or grazing(i, a[i]))) or \
                (b[i] >= 0 and (near_edge(i, b[i]) or grazing(i, b[i]))):
            adj += 1
            continue
        if a[i] < 0 or b[i] < 0:
            continue
        if set(I[a[i]].tolist()) & set(I[b[i]].tolist()):
            adj += 1
        elif abs(wt_hits[i, 0] - mt_hits[i, 0]) <= 1e-5 * abs(wt_hits[i, 0]):
            adj += 1
    ok = (~mism) & (a >= 0)
    pr = a[ok]
    A, B, C = P[I[pr, 0]], P[I[pr, 1]], P[I[pr, 2]]
    n = np.cross(B - A, C - A)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = rays[ok, 4:7].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    cos = np.abs((n * d).sum(1))
    rel = np.abs(wt_hits[ok, 0] - mt_hits[ok, 0]) / np.abs(mt_hits[ok, 0])
    k = cos >= min_cos
    return mism.mean(), (adj / max(1, mism.sum()) if mism.any() else 1.0), (rel[k].max() if k.any() else 0.0), k.mean()
