"""Several light samples per vertex (PTRendererParams::direct_light_sample_num,
src/factory/renderer.hpp:18): the reference averages that many rounds of
light sampling and BSDF sampling at every surface and medium vertex
(src/renderer/path_trace_renderer.cpp:100-106, 143-151).  The GPU draws the
light samples and keeps its one continuation ray as the BSDF sample, with the
multi-sample power heuristic (n_light = direct_light_sample_num, n_bsdf = 1),
which has the same expectation:

* configuration 1 against the oracle in reference mode and configuration 3
  (fog) in fixed mode, both with 3 light samples per vertex;
* the closed emitting box keeps its closed form;
* more light samples per vertex lower the error of direct lighting at equal
  samples per pixel;
* the films are reproducible (contributions are added in sample order).
"""
import numpy as np
import pytest

from tests.furnace import furnace_box, furnace_expected
from tests.test_gpu_render import _compare_estimators

pytestmark = pytest.mark.gpu


def test_c1_three_light_samples_match_reference_estimator(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    _compare_estimators(gpu_ctx, s, 32, 32, 256, 32, 8, 8, quirks=True, nee=0, rel_max=0.01,
                        direct_light_sample_num=3)


def test_c3_fog_two_light_samples_match_fixed_estimator(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(3, detail=1 / 12)
    _compare_estimators(gpu_ctx, s, 32, 32, 256, 32, 8, 8, quirks=False, nee=0, max_depth=12,
                        direct_light_sample_num=2)


@pytest.mark.parametrize("nls", [2, 4])
def test_furnace_with_several_light_samples(gpu_ctx, nls):
    rho = (0.5, 0.6, 0.7)
    s = furnace_box(rho=rho)
    gpu_ctx.load_scene(s)
    sums, _, st = gpu_ctx.render(s.render_params(width=64, height=64, spp=128, seed=4, max_depth=6, min_depth=1,
                                                 direct_light_sample_num=nls), want_image=False)
    img = sums[..., :3] / sums[..., 3:]
    px = img.reshape(-1, 3)
    z = (px.mean(0) - furnace_expected((1, 1, 1), rho, 6)) / (px.std(0, ddof=1) / np.sqrt(px.shape[0]))
    print(f"furnace with {nls} light samples: z {z}, shadow rays per extension ray "
          f"{st.shadow_rays / st.extension_rays:.2f}")
    assert np.abs(z).max() < 5.0
    assert st.shadow_rays > 0.6 * nls * st.extension_rays


def test_more_light_samples_lower_the_error(gpu_ctx):
    """Direct light only (max depth 1) on configuration 4 with 4 spheres, whose
    light is mostly hidden behind an aperture: light sampling is the noise."""
    from tracer_amd import native
    s = native.config_scene(4, detail=1 / 16)
    gpu_ctx.load_scene(s)
    W = H = 64
    R = None
    for c in range(8):
        R, _, _ = gpu_ctx.render(s.render_params(width=W, height=H, spp=4096, seed=1, sample_offset=4096 * c,
                                                 max_depth=1), film=R, want_image=False)
    R = R[..., :3] / R[..., 3:]

    def err(nls, seed):
        sums, _, _ = gpu_ctx.render(s.render_params(width=W, height=H, spp=32, seed=seed, max_depth=1,
                                                    direct_light_sample_num=nls), want_image=False)
        img = sums[..., :3] / sums[..., 3:]
        return float(np.mean((img - R) ** 2 / (R ** 2 + 1e-2)))

    e1 = np.median([err(1, 20 + k) for k in range(4)])
    e4 = np.median([err(4, 30 + k) for k in range(4)])
    print(f"relMSE at 32 spp: 1 light sample {e1:.5f}, 4 light samples {e4:.5f}")
    assert e4 < 0.5 * e1


def test_several_light_samples_are_reproducible_and_checked(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    p = s.render_params(width=32, height=32, spp=8, seed=3, direct_light_sample_num=3)
    a, _, _ = gpu_ctx.render(p)
    b, _, _ = gpu_ctx.render(p)
    assert np.array_equal(a, b)
    for kw in ({"direct_light_sample_num": 3, "guide": 1}, {"direct_light_sample_num": 65},
               {"direct_light_sample_num": -1}):
        with pytest.raises(native.TracerError):
            gpu_ctx.render(s.render_params(width=8, height=8, spp=1, seed=1, **kw), want_image=False)
