"""Parity at the benchmark configurations' real sizes.

* Traversal (BVHAccel::intersect_p / intersect, src/accelerator/bvh.cpp:112-185,
  with the watertight triangle test): on configuration 2 (1,310,754
  triangles) 2^20 camera rays, 2^20 random rays, 2^20 shadow rays from hit
  points to points on the lights and 2^18 rays aimed at triangle interiors;
  on configuration 5 (18,874,880 triangles) 2^18 of each kind.  Closest hits
  (primitive id, t and barycentrics) and any-hit answers equal the oracle's
  traversal of its own binary SAH BVH bit for bit; on the interior rays the
  reference's Moller-Trumbore test agrees with t within 1e-5 relative on the
  conditioned domain (triangle.cpp:145-175).
* Images (PathTraceRenderer::eval_pixel_li, path_trace_renderer.cpp:27-210):
  configuration 4 with its full geometry and configuration 5 (fog, sun, 512
  lights) against the oracle in fixed mode on small films (tile and image
  z-scores); on configuration 1 the GPU and the oracle converge to one
  65536-spp GPU reference pixel by pixel: relMSE falls as 1 / spp for both,
  and the relMSE between the two images at matched spp is the sum of their
  errors against the reference (no bias between them).
"""
import numpy as np
import pytest

from tests.raygen import compare_with_mt, hit_prims, interior_rays, random_rays
from tests.test_gpu_render import _seeds_tiles

pytestmark = pytest.mark.gpu


def _check_exact(gh, oh, what):
    gp, op = hit_prims(gh), hit_prims(oh)
    assert np.array_equal(gp, op), f"{what}: {np.count_nonzero(gp != op)} prim mismatches"
    hit = gp >= 0
    assert np.array_equal(gh[hit, 0], oh[hit, 0]), f"{what}: t differs"
    assert np.array_equal(gh[hit, 2:], oh[hit, 2:]), f"{what}: barycentrics differ"
    return hit.mean()


def _camera_batch(ctx, scene, n, seed):
    W, H = scene.defaults.width, scene.defaults.height
    rng = np.random.default_rng(seed)
    pix = rng.integers(0, W * H, n)
    film = np.stack([(pix % W + rng.uniform(0, 1, n)) / W, (pix // W + rng.uniform(0, 1, n)) / H], 1)
    return ctx.camera_rays(film.astype(np.float32), rng.uniform(0, 1, (n, 2)).astype(np.float32))


def _shadow_batch(scene, hits, rays, n, seed):
    """Segments from hit points (offset along the ray) to random points on random emissive triangles."""
    rng = np.random.default_rng(seed)
    ok = np.nonzero(hit_prims(hits) >= 0)[0]
    pick = ok[rng.integers(0, len(ok), n)]
    p = rays[pick, :3].astype(np.float64) + rays[pick, 4:7] * (hits[pick, 0:1] * 0.999)
    P = scene.positions.reshape(-1, 3).astype(np.float64)
    I = scene.indices.reshape(-1, 3)
    em = np.nonzero((scene.emission.reshape(-1, 3) != 0).any(1))[0]
    tri = em[rng.integers(0, len(em), n)]
    w = rng.dirichlet((1, 1, 1), n)
    q = P[I[tri, 0]] * w[:, :1] + P[I[tri, 1]] * w[:, 1:2] + P[I[tri, 2]] * w[:, 2:3]
    d = q - p
    dist = np.linalg.norm(d, axis=1)
    out = np.zeros((n, 8), np.float32)
    out[:, :3] = p
    out[:, 3] = 1e-4
    out[:, 4:7] = d / dist[:, None]
    out[:, 7] = dist * (1 - 1e-4)
    return out


def _traversal_parity(ctx, scene, n, n_interior):
    from oracle import refcpu
    ref = refcpu.RefScene(scene, 3)
    cam = _camera_batch(ctx, scene, n, 1)
    rnd = random_rays(scene, n, seed=2)
    inter, _ = interior_rays(scene, n_interior, seed=3)
    print(f"{scene.num_triangles} triangles", flush=True)
    for name, rays in (("camera", cam), ("random", rnd), ("interior", inter)):
        gh = ctx.trace_closest(rays)
        frac = _check_exact(gh, ref.closest_wt(rays, threads=16), name)
        print(f"  {name}: {len(rays)} rays bit-exact, {frac:.2f} hit", flush=True)
        if name == "interior":
            r = compare_with_mt(scene, rays, gh, ref.closest_mt(rays, threads=16))
            print(f"  interior vs reference MT: {r}", flush=True)
            assert r["well_fraction"] > 0.8 and r["well_id_mismatch"] == 0 and r["well_t_rel_max"] <= 1e-5
        if name == "random":
            sh = _shadow_batch(scene, gh, rays, n, 4)
            og, oo = ctx.trace_any(sh), ref.any_wt(sh, threads=16)
            assert np.array_equal(og, oo), f"shadow: {np.count_nonzero(og != oo)} mismatches"
            print(f"  shadow: {len(sh)} rays equal, {og.mean():.2f} occluded", flush=True)


def test_traversal_parity_configuration_2_full(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(2)
    gpu_ctx.load_scene(s)
    _traversal_parity(gpu_ctx, s, 1 << 20, 1 << 18)


@pytest.fixture(scope="module")
def c5_scene():
    from tracer_amd import native
    return native.config_scene(5)


def test_traversal_parity_configuration_5_full(gpu_ctx, c5_scene):
    gpu_ctx.load_scene(c5_scene)
    _traversal_parity(gpu_ctx, c5_scene, 1 << 18, 1 << 16)


def _compare_images(ctx, scene, W, H, gpu_spp, cpu_spp, seeds, tile, nee, z_max=5.0, **kw):
    from oracle import refcpu
    ctx.load_scene(scene)
    ref = refcpu.RefScene(scene, 3)

    def gpu(seed):
        return ctx.render(scene.render_params(width=W, height=H, spp=gpu_spp, seed=seed, **kw), want_image=False)[0]

    def cpu(seed):
        return ref.render(scene.render_params(width=W, height=H, spp=cpu_spp, seed=9000 + seed, **kw), quirks=False,
                          nee=nee, threads=16)[0]

    g = _seeds_tiles(gpu, range(seeds), tile)
    c = _seeds_tiles(cpu, range(seeds), tile)
    se = np.sqrt(g.var(0, ddof=1) / seeds + c.var(0, ddof=1) / seeds)
    z = (g.mean(0) - c.mean(0)) / np.maximum(se, 1e-6)
    ig, ic = g.mean(axis=(1, 2, 3)), c.mean(axis=(1, 2, 3))
    zi = (ig.mean() - ic.mean()) / np.sqrt(ig.var(ddof=1) / seeds + ic.var(ddof=1) / seeds)
    print(f"max tile |z| {np.abs(z).max():.2f}, image z {zi:.2f} (gpu {ig.mean():.4f}, cpu {ic.mean():.4f})")
    assert np.abs(z).max() < z_max and abs(zi) < z_max


def test_configuration_4_full_geometry_matches_oracle(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(4)
    _compare_images(gpu_ctx, s, 32, 32, 1024, 128, 8, 8, nee=0, max_depth=8)


def test_configuration_5_full_matches_oracle(gpu_ctx, c5_scene):
    """Fog, the sun, the environment and 512 quad lights; the oracle samples
    one light by power (its all-lights mode would trace 514 shadow rays per
    vertex)."""
    _compare_images(gpu_ctx, c5_scene, 32, 32, 256, 32, 8, 8, nee=1, max_depth=6)


def test_configuration_1_converges_pixel_by_pixel(gpu_ctx):
    from oracle import refcpu
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    ref = refcpu.RefScene(s, 3)
    W = H = 32
    R = None
    for c in range(8):
        R, _, _ = gpu_ctx.render(s.render_params(width=W, height=H, spp=8192, seed=1, sample_offset=8192 * c), film=R,
                                 want_image=False)
    Rimg = R[..., :3] / R[..., 3:]

    def img(sums):
        return sums[..., :3] / sums[..., 3:]

    def relmse(a, b):
        return float(np.mean((a - b) ** 2 / (b ** 2 + 1e-2)))

    spps = (256, 1024, 4096)
    eg, ec, imgs = [], [], {}
    for n in spps:
        # three seeds per spp: one image's relMSE moves with its brightest outliers
        gs = [img(gpu_ctx.render(s.render_params(width=W, height=H, spp=n, seed=40 + n + 7 * k), want_image=False)[0])
              for k in range(3)]
        cs = [img(ref.render(s.render_params(width=W, height=H, spp=n, seed=50 + n + 7 * k), quirks=False, nee=0,
                             threads=16)[0]) for k in range(3)]
        eg.append(float(np.mean([relmse(g, Rimg) for g in gs])))
        ec.append(float(np.mean([relmse(c, Rimg) for c in cs])))
        imgs[n] = (gs, cs)
    floor = eg[-1] * spps[-1] / 65536.0   # the reference's own error, from the GPU's 1 / spp law
    sg = np.polyfit(np.log(spps), np.log(np.array(eg) - floor), 1)[0]
    sc = np.polyfit(np.log(spps), np.log(np.array(ec) - floor), 1)[0]
    between = float(np.mean([relmse(g, c) for g, c in zip(*imgs[4096])]))
    print(f"relMSE vs 65536-spp reference: GPU {eg}, oracle {ec}; slopes {sg:.3f}, {sc:.3f}; "
          f"GPU vs oracle at 4096 spp {between:.3e} (sum of errors {eg[-1] + ec[-1]:.3e})")
    assert -1.15 < sg < -0.85 and -1.15 < sc < -0.85
    # independent unbiased images: E|g - c|^2 = E|g - R|^2 + E|c - R|^2 (up to the reference's own error)
    assert 0.75 < between / (eg[-1] + ec[-1] - 2 * floor) < 1.3
