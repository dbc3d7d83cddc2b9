"""Pins for the oracle's ray/triangle tests and BVHs, and invariants of the
product's host-built 4-wide BVH.

* Known answers with analytic t and barycentrics (edge, vertex, parallel,
  back-face cases) for the reference's Moller-Trumbore restatement
  (src/shape/triangle.cpp:145-175) and the watertight test.
* The oracle's BVH traversals (reference BVH + MT; conservative traversal +
  watertight) equal brute force over all triangles.
* Reference BVH invariants (src/accelerator/bvh.cpp:191-306): every primitive
  in exactly one leaf, children inside parents, leaves within max_leaf unless
  all centroids coincide.
* 4-wide BVH invariants of tr_bvh4_export (the layout the GPU walks).
"""
import ctypes as C

import numpy as np
import pytest

from tests.raygen import hit_prims, interior_rays, random_rays

TRI = [0, 0, 0, 1, 0, 0, 0, 1, 0]


def ray(o, d, tmin=0.0, tmax=1e30):
    return list(o) + [tmin] + list(d) + [tmax]


@pytest.mark.parametrize("fn_name", ["mt_single", "wt_single"])
def test_triangle_known_answers(fn_name):
    from oracle import refcpu
    fn = getattr(refcpu, fn_name)
    ok, out = fn(ray((0.25, 0.25, 1), (0, 0, -1)), TRI)
    assert ok and out[0] == 1.0 and out[1] == 0.25 and out[2] == 0.25
    ok, out = fn(ray((0.2, 0.5, -2), (0, 0, 1)), TRI)   # back side: both tests are double-sided
    assert ok and out[0] == 2.0 and abs(out[1] - 0.2) < 1e-7 and abs(out[2] - 0.5) < 1e-7
    assert fn(ray((0.5, 0.0, 1), (0, 0, -1)), TRI)[0]      # on edge AB (inclusive)
    assert fn(ray((0.0, 0.0, 1), (0, 0, -1)), TRI)[0]      # on vertex A
    assert fn(ray((0.5, 0.5, 1), (0, 0, -1)), TRI)[0]      # on edge BC
    assert not fn(ray((0.6, 0.6, 1), (0, 0, -1)), TRI)[0]  # outside
    assert not fn(ray((0.2, 0.2, 1), (1, 0, 0)), TRI)[0]   # parallel to the plane
    assert not fn(ray((0.25, 0.25, 1), (0, 0, -1), tmax=0.5), TRI)[0]
    assert not fn(ray((0.25, 0.25, 1), (0, 0, -1), tmin=1.5), TRI)[0]


def test_watertight_on_shared_edges():
    """Rays aimed at the diagonal shared by the two triangles of an exactly
    planar quad (z = 0, all coordinates representable) cross the plane inside
    the quad, so the watertight test must report a hit on one of them for
    every ray; the reference's Moller-Trumbore is allowed to leak."""
    from oracle import refcpu
    rng = np.random.default_rng(0)
    t1 = np.array([0, 0, 0, 1, 0, 0, 1, 1, 0], np.float32)
    t2 = np.array([0, 0, 0, 1, 1, 0, 0, 1, 0], np.float32)
    wt_miss = mt_miss = 0
    for _ in range(4000):
        sx = np.float32(rng.uniform(0.01, 0.99))
        p = np.array([sx, sx, 0], np.float32)
        o = (p + np.array([rng.normal(), rng.normal(), abs(rng.normal()) + 0.1])).astype(np.float32)
        r = ray(o, (p - o).astype(np.float32))
        wt_miss += not (refcpu.wt_single(r, t1)[0] or refcpu.wt_single(r, t2)[0])
        mt_miss += not (refcpu.mt_single(r, t1)[0] or refcpu.mt_single(r, t2)[0])
    print(f"shared-edge misses: watertight {wt_miss}, reference MT {mt_miss} of 4000")
    assert wt_miss == 0


@pytest.fixture(scope="module")
def c1():
    from tracer_amd import native
    from oracle import refcpu
    s = native.config_scene(1)
    return s, refcpu.RefScene(s, 3)


@pytest.fixture(scope="module")
def c2_small():
    from tracer_amd import native
    from oracle import refcpu
    s = native.config_scene(2, detail=1 / 64)   # one sphere, ground, 16 lights
    return s, refcpu.RefScene(s, 3)


def test_oracle_traversals_equal_brute_force(c1, c2_small):
    for s, ref in (c1, c2_small):
        rays = random_rays(s, 4096, seed=11)
        b = ref.closest_brute(rays)
        w = ref.closest_wt(rays)
        assert np.array_equal(b.view(np.int32), w.view(np.int32))
        m = ref.closest_mt(rays)
        assert (hit_prims(m) != hit_prims(b)).mean() < 2e-3
        assert not np.any(hit_prims(b) >= s.num_triangles)
        rays[:, 7] = np.random.default_rng(3).uniform(0.01, 5.0, rays.shape[0])
        occ = ref.any_wt(rays)
        bb = ref.closest_brute(rays)
        assert np.array_equal(occ, (hit_prims(bb) >= 0).astype(np.int32))


def test_interior_rays_mt_vs_watertight(c2_small):
    """Rays aimed at triangle interiors: on the well-conditioned suite the
    watertight closest hit and the reference's MT closest hit name the same
    triangle and agree in t within 1e-5 relative (tests.raygen.compare_with_mt)."""
    from tests.raygen import compare_with_mt
    s, ref = c2_small
    rays, _ = interior_rays(s, 1 << 14, seed=12)
    b, m = ref.closest_brute(rays), ref.closest_mt(rays)
    r = compare_with_mt(s, rays, b, m)
    print(r)
    assert r["well_fraction"] > 0.9 and r["well_id_mismatch"] == 0 and r["well_t_rel_max"] <= 1e-5
    assert r["all_id_mismatch_rate"] < 2e-3


def test_reference_bvh_invariants(c2_small):
    s, ref = c2_small
    nodes, prims = ref.bvh()
    assert sorted(prims.tolist()) == list(range(s.num_triangles))
    P = s.positions.reshape(-1, 3)
    I = s.indices.reshape(-1, 3)
    seen = np.zeros(s.num_triangles, np.int32)

    def walk(i, lo, hi):
        n = nodes[i]
        assert np.all(n["lo"] >= lo) and np.all(n["hi"] <= hi)
        if n["count"] > 0:
            ps = prims[n["offset"]:n["offset"] + n["count"]]
            seen[ps] += 1
            v = P[I[ps].reshape(-1)]
            assert np.all(v >= n["lo"]) and np.all(v <= n["hi"])
            if n["count"] > 3:
                c = (P[I[ps]].min(axis=1) + P[I[ps]].max(axis=1)) * 0.5
                assert np.all(c.max(axis=0) == c.min(axis=0))
            return
        walk(i + 1, n["lo"], n["hi"])
        walk(n["offset"], n["lo"], n["hi"])

    import sys
    sys.setrecursionlimit(10000)
    walk(0, nodes[0]["lo"], nodes[0]["hi"])
    assert np.all(seen == 1)


def test_bvh4_layout_invariants(c2_small):
    from tracer_amd import native, abi
    s, _ = c2_small
    L = native.hip_lib()
    L.tr_bvh4_export.argtypes = [C.POINTER(abi.TrSceneDesc), C.POINTER(abi.TrBvhOpts), C.c_void_p, C.c_int64, C.c_void_p,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    d = s.desc()
    o = abi.TrBvhOpts(4, 0)
    nn, depth = C.c_int64(), C.c_int32()
    assert L.tr_bvh4_export(C.byref(d), C.byref(o), None, 0, None, C.byref(nn), C.byref(depth)) == 0
    nodes = np.zeros((nn.value, 32), np.float32)
    tris = np.zeros((s.num_triangles, 12), np.float32)
    assert L.tr_bvh4_export(C.byref(d), C.byref(o), nodes.ctypes.data, nn.value, tris.ctypes.data, C.byref(nn),
                            C.byref(depth)) == 0
    child = nodes[:, 24:28].copy().view(np.int32)
    lo = np.stack([nodes[:, 0:4], nodes[:, 8:12], nodes[:, 16:20]], axis=-1)
    hi = np.stack([nodes[:, 4:8], nodes[:, 12:16], nodes[:, 20:24]], axis=-1)
    prim = tris[:, 3].copy().view(np.int32)
    assert sorted(prim.tolist()) == list(range(s.num_triangles))
    covered = np.zeros(s.num_triangles, np.int32)
    P = s.positions.reshape(-1, 3)
    I = s.indices.reshape(-1, 3)
    parent_of = {}
    for n in range(nn.value):
        for c in range(4):
            ch = child[n, c]
            if ch >= 0:
                assert ch not in parent_of
                parent_of[ch] = (n, c)
                clo = lo[ch][np.isfinite(lo[ch][:, 0])]
                chi = hi[ch][np.isfinite(hi[ch][:, 0])]
                assert np.all(clo >= lo[n, c]) and np.all(chi <= hi[n, c])
            elif ch != -1:
                code = ~ch
                first, cnt = code >> 4, code & 15
                assert 1 <= cnt <= 4
                covered[first:first + cnt] += 1
                v = tris[first:first + cnt][:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3)
                assert np.all(v >= lo[n, c]) and np.all(v <= hi[n, c])
                # packed vertices are the original triangle's vertices
                pv = P[I[prim[first:first + cnt]]].reshape(-1, 3)
                assert np.array_equal(v, pv.astype(np.float32))
    assert np.all(covered == 1)
    assert len(parent_of) == nn.value - 1 and 0 not in parent_of
