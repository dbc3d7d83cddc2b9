"""Closest-hit / any-hit parity of the GPU traversal (tr_trace_closest /
tr_trace_any) against the oracle: a brute-force loop over all triangles with
the same watertight test (bit-exact prim id and t), and the reference's
Moller-Trumbore + binary BVH (src/accelerator/bvh.cpp:149-185,
src/shape/triangle.cpp:145-175): prim id equal and t within 1e-5 relative on
rays aimed at triangle interiors."""
import numpy as np
import pytest

from tests.raygen import compare_with_mt, edge_case_rays, hit_prims, interior_rays, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(gpu_ctx):
    from tracer_amd import native
    from oracle import refcpu
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    return s, refcpu.RefScene(s, 3)


def _check_exact(gh, oh):
    gp, op = hit_prims(gh), hit_prims(oh)
    assert np.array_equal(gp, op), f"{np.count_nonzero(gp != op)} prim mismatches"
    hit = gp >= 0
    assert np.array_equal(gh[hit, 0], oh[hit, 0]), "t differs"
    assert np.array_equal(gh[hit, 2:], oh[hit, 2:]), "barycentrics differ"


def test_c1_random_rays_bit_exact(gpu_ctx, c1):
    s, ref = c1
    gpu_ctx.load_scene(s)
    rays = random_rays(s, 1 << 16, seed=1)
    _check_exact(gpu_ctx.trace_closest(rays), ref.closest_brute(rays))


def test_c1_interior_rays_vs_reference_mt(gpu_ctx, c1):
    s, ref = c1
    gpu_ctx.load_scene(s)
    rays, _ = interior_rays(s, 1 << 15, seed=2)
    gh = gpu_ctx.trace_closest(rays)
    _check_exact(gh, ref.closest_brute(rays))
    r = compare_with_mt(s, rays, gh, ref.closest_mt(rays))
    print("c1 GPU vs reference MT:", r)
    assert r["well_fraction"] > 0.9 and r["well_id_mismatch"] == 0 and r["well_t_rel_max"] <= 1e-5


def test_c1_any_hit(gpu_ctx, c1):
    s, ref = c1
    gpu_ctx.load_scene(s)
    rays = random_rays(s, 1 << 15, seed=3)
    rays[:, 7] = np.random.default_rng(4).uniform(0.01, 3.0, rays.shape[0])
    assert np.array_equal(gpu_ctx.trace_any(rays), ref.any_wt(rays))


def test_c1_edge_case_rays_bit_exact(gpu_ctx, c1):
    """Rays at shared vertices and edges, axis-aligned rays, distant origins and
    origins on the surface: the widened box test must cull no box holding the
    closest hit, so hits equal the brute-force loop exactly."""
    s, ref = c1
    gpu_ctx.load_scene(s)
    rays = edge_case_rays(s, 5 * 4096, seed=7)
    assert np.any(rays[:, 4:7] == 0.0), "expected axis-aligned rays"
    _check_exact(gpu_ctx.trace_closest(rays), ref.closest_brute(rays))
    short = rays.copy()
    short[:, 7] = np.random.default_rng(8).uniform(0.01, 3.0, rays.shape[0])
    assert np.array_equal(gpu_ctx.trace_any(short), ref.any_wt(short))


@pytest.mark.parametrize("detail", [0.125])
def test_c2_subset_bit_exact(gpu_ctx, detail):
    from tracer_amd import native
    from oracle import refcpu
    s = native.config_scene(2, detail=detail)
    gpu_ctx.load_scene(s)
    ref = refcpu.RefScene(s, 3)
    rays = random_rays(s, 1 << 14, seed=5)
    gh = gpu_ctx.trace_closest(rays)
    _check_exact(gh, ref.closest_wt(rays))
    rays2, _ = interior_rays(s, 1 << 14, seed=6)
    gh2 = gpu_ctx.trace_closest(rays2)
    _check_exact(gh2, ref.closest_wt(rays2))
    r = compare_with_mt(s, rays2, gh2, ref.closest_mt(rays2))
    print("c2 subset GPU vs reference MT:", r)
    assert r["well_fraction"] > 0.9 and r["well_id_mismatch"] == 0 and r["well_t_rel_max"] <= 1e-5
    assert r["all_id_mismatch_rate"] < 2e-3
