"""Learned light sampling (tr_render_params::guide_light_grid, guide = 2 and
guide = 1): next-event estimation draws lights and parts of lights from a
per-cell distribution trained on the renderer's own unoccluded light samples,
mixed with the power-based choice.  The estimator keeps the expectation of
PathTraceRenderer::eval_pixel_li (src/renderer/path_trace_renderer.cpp:27-210):

* the closed emitting box still gives Le * sum_{k<=D} rho^k;
* per-tile means equal the plain renderer's on configurations 4 (one light
  behind an aperture), 2 (16 quad lights and the constant environment) and 5
  (many lights, the sun, fog), for guide = 2 and with the network (guide = 1);
* on configuration 4 it lowers the error at equal samples;
* training sums are integers, so a render is reproducible, and a render that
  trains from sample 0 starts from the power-based choice again.
"""
import numpy as np
import pytest

from tests.furnace import furnace_box, furnace_expected
from tests.test_gpu_render import _seeds_tiles

pytestmark = pytest.mark.gpu


def _lc_render(ctx, scene, W, H, spp, seed, grid, guide=2, train_spp=4, train_passes=1, **kw):
    """Training passes of train_spp each, then the rest sampling the cache."""
    ts = train_spp * train_passes
    s1, _, a = ctx.render(scene.render_params(width=W, height=H, spp=ts, seed=seed, guide=guide,
                                              guide_train_passes=train_passes, samples_per_pass=train_spp,
                                              guide_light_grid=grid, **kw), want_image=False)
    s2, _, b = ctx.render(scene.render_params(width=W, height=H, spp=spp - ts, seed=seed, sample_offset=ts, guide=guide,
                                              guide_train_passes=0, guide_light_grid=grid, **kw),
                          film=s1, want_image=False)
    assert a.nonfinite_dropped == 0 and b.nonfinite_dropped == 0
    return s2


@pytest.mark.parametrize("max_depth,min_depth", [(4, 10), (6, 1)])
def test_light_cache_furnace_closed_form(gpu_ctx, max_depth, min_depth):
    rho = (0.5, 0.6, 0.7)
    s = furnace_box(rho=rho)
    gpu_ctx.load_scene(s)
    sums = _lc_render(gpu_ctx, s, 64, 64, 128, 5, 4, max_depth=max_depth, min_depth=min_depth)
    img = sums[..., :3] / sums[..., 3:]
    px = img.reshape(-1, 3)
    m, se = px.mean(0), px.std(0, ddof=1) / np.sqrt(px.shape[0])
    z = (m - furnace_expected((1, 1, 1), rho, max_depth)) / se
    print(f"light-cache furnace D={max_depth}: mean {m}, z {z}")
    assert np.abs(z).max() < 5.0


def _compare_to_plain(ctx, scene, W, H, spp, seeds, tile, grid, guide, z_max=5.0, **kw):
    """Per-tile and whole-image means of independent seeds: the cached
    renderer against the plain one, as z-scores."""
    ctx.load_scene(scene)

    def plain(seed):
        sums, _, _ = ctx.render(scene.render_params(width=W, height=H, spp=spp, seed=seed, **kw), want_image=False)
        return sums

    def cached(seed):
        if guide == 1:
            ctx.guide_init(hidden=64, seed=3 + seed, batch=1 << 14, steps_per_pass=8)
        return _lc_render(ctx, scene, W, H, spp, 7000 + seed, grid, guide=guide, **kw)

    g = _seeds_tiles(cached, range(seeds), tile)
    c = _seeds_tiles(plain, range(seeds), tile)
    mg, mc = g.mean(0), c.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / g.shape[0] + c.var(0, ddof=1) / c.shape[0])
    z = (mg - mc) / np.maximum(se, 1e-6)
    ig, ic = g.mean(axis=(1, 2, 3)), c.mean(axis=(1, 2, 3))
    zi = (ig.mean() - ic.mean()) / np.sqrt(ig.var(ddof=1) / len(ig) + ic.var(ddof=1) / len(ic))
    print(f"light cache vs plain: max tile |z| {np.abs(z).max():.2f}, image z {zi:.2f} "
          f"(means {ig.mean():.4f} vs {ic.mean():.4f})")
    assert np.abs(z).max() < z_max
    assert abs(zi) < z_max


@pytest.mark.parametrize("guide", [2, 1])
def test_light_cache_matches_plain_on_occluded_light(gpu_ctx, guide):
    from tracer_amd import native
    s = native.config_scene(4, detail=1 / 16)
    _compare_to_plain(gpu_ctx, s, 32, 32, 256, 8, 8, 8, guide, max_depth=8)


def test_light_cache_matches_plain_with_environment(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(2, detail=1 / 64)
    s.camera.pos[:] = (0.0, 2.0, 14.0)
    _compare_to_plain(gpu_ctx, s, 32, 32, 128, 8, 8, 6, 2, max_depth=6)


def test_light_cache_matches_plain_with_sun_and_fog(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(5, detail=1 / 2048)
    _compare_to_plain(gpu_ctx, s, 32, 32, 128, 8, 8, 8, 2, max_depth=6)


def test_light_cache_lowers_error_on_occluded_light(gpu_ctx):
    """Configuration 4 with 4 spheres at 128 x 128, 256 spp: the learned light
    sampling (4 training passes of 8 spp, 16^3 cells) against plain at equal
    samples, both against a 32768-spp plain reference."""
    from tracer_amd import native
    s = native.config_scene(4, detail=1 / 16)
    gpu_ctx.load_scene(s)
    W = H = 128
    ref = None
    for c in range(16):
        ref, _, _ = gpu_ctx.render(s.render_params(width=W, height=H, spp=2048, seed=1, sample_offset=2048 * c),
                                   film=ref, want_image=False)
    R = ref[..., :3] / ref[..., 3:]

    def relmse(sums):
        img = sums[..., :3] / sums[..., 3:]
        return float(np.mean((img - R) ** 2 / (R ** 2 + 1e-2)))

    plain = [relmse(gpu_ctx.render(s.render_params(width=W, height=H, spp=256, seed=50 + k), want_image=False)[0])
             for k in range(6)]
    cached = [relmse(_lc_render(gpu_ctx, s, W, H, 256, 60 + k, 16, train_spp=8, train_passes=4)) for k in range(6)]
    gain = np.median(plain) / np.median(cached)
    print(f"equal-spp relMSE: plain {np.median(plain):.5f}, light cache {np.median(cached):.5f}, gain {gain:.2f}")
    # measured 1.48 to 1.53 over builds whose sampling differs only in its random realisation
    assert gain > 1.3


def test_light_cache_is_reproducible_and_restarts(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(4, detail=1 / 16)
    other = native.config_scene(1)
    gpu_ctx.load_scene(s)
    a = _lc_render(gpu_ctx, s, 32, 32, 32, 9, 8, train_passes=2)
    b = _lc_render(gpu_ctx, s, 32, 32, 32, 9, 8, train_passes=2)   # trains from sample 0 again: same film
    assert np.array_equal(a, b)
    gpu_ctx.load_scene(other)
    gpu_ctx.render(other.render_params(width=16, height=16, spp=4, seed=1, guide=2, guide_train_passes=1,
                                       guide_light_grid=4), want_image=False)
    gpu_ctx.load_scene(s)
    c = _lc_render(gpu_ctx, s, 32, 32, 32, 9, 8, train_passes=2)
    assert np.array_equal(a, c)


def test_light_cache_arguments(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    for kw in ({"guide": 2, "guide_light_grid": 0}, {"guide": 2, "guide_light_grid": 65},
               {"guide": 3, "guide_light_grid": 4}):
        with pytest.raises(native.TracerError):
            gpu_ctx.render(s.render_params(width=8, height=8, spp=1, seed=1, **kw), want_image=False)
