"""Closed-form checks of the parts with no reference counterpart: the
directional sun light and the box filter.  A Lambert plane lit only by the sun
has radiance rho / pi * E * cos(theta) at every point, and with one delta light
and no further bounce every sample of a pixel has exactly that value."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO = np.array([0.5, 0.3, 0.2])
SUN_DIR = np.array([0.3, 1.0, 0.2])
SUN_E = np.array([3.0, 2.0, 1.0])


def _sun_plane_scene():
    from tracer_amd import abi, native
    P = np.array([[-10, 0, -10], [10, 0, -10], [10, 0, 10], [-10, 0, 10]], np.float32)
    idx = np.array([[0, 2, 1], [0, 3, 2]], np.int32)   # wound so the geometric normal is +y
    n = np.cross(P[2] - P[0], P[1] - P[0])
    assert n[1] > 0
    mat = abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=tuple(RHO))
    cam = abi.TrCamera()
    cam.pos = abi.f3(0.0, 5.0, 0.001)
    cam.target = abi.f3(0.0, 0.0, 0.0)
    cam.up = abi.f3(0.0, 0.0, 1.0)
    cam.fov_y = np.radians(40.0)
    cam.lens_radius = 0.0
    cam.focal_distance = 1.0
    cam.aspect = 1.0
    s = native.Scene(positions=P, indices=idx, material_ids=np.zeros(2, np.int32), materials=[mat],
                     sun=(tuple(SUN_DIR / np.linalg.norm(SUN_DIR)), tuple(SUN_E)), camera=cam)
    return s


def test_sun_irradiance_on_a_plane(gpu_ctx):
    s = _sun_plane_scene()
    gpu_ctx.load_scene(s)
    p = s.render_params(width=32, height=32, spp=4, seed=3, max_depth=1, min_depth=1)
    sums, img, st = gpu_ctx.render(p)
    cos_t = SUN_DIR[1] / np.linalg.norm(SUN_DIR)
    expect = RHO / np.pi * SUN_E * cos_t
    rel = np.abs(img - expect) / expect
    print(f"sun on a plane: expected {expect}, max rel error {rel.max():.2e}")
    assert rel.max() < 1e-5
    assert st.shadow_rays > 0


def test_box_filter_weights(gpu_ctx):
    from tracer_amd import abi
    s = _sun_plane_scene()
    gpu_ctx.load_scene(s)
    spp = 16
    # radius 0.5: each sample lands in its own pixel with weight 1
    p = s.render_params(width=32, height=32, spp=spp, seed=4, max_depth=1, min_depth=1,
                        filter_type=abi.TR_FILTER_BOX, filter_radius=0.5)
    sums, img, _ = gpu_ctx.render(p)
    assert np.all(sums[..., 3] == spp)
    # radius 1: samples of the 3x3 neighbourhood within the box; the image of a
    # constant scene stays constant and an interior pixel gathers 4 spp on average
    p = s.render_params(width=32, height=32, spp=spp, seed=4, max_depth=1, min_depth=1,
                        filter_type=abi.TR_FILTER_BOX, filter_radius=1.0)
    sums, img1, _ = gpu_ctx.render(p)
    assert np.allclose(img1, img, rtol=1e-5)
    w = sums[1:-1, 1:-1, 3]
    print(f"box radius 1: mean interior weight {w.mean():.3f} (expected {4 * spp})")
    assert abs(w.mean() - 4 * spp) < 0.02 * 4 * spp
