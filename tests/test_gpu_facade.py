"""The drop-in path through the C++ facade: a compiled program
(tests/cpp/facade_harness.cpp) builds configuration 1 with the tracer::
factories and calls create_pt_renderer(params)->render(scene,
Film(res, create_gaussin_filter(0.5, 0.6))) as a caller of the reference
does (src/core/renderer.hpp:16, src/main.cpp).  Its images equal the C ABI's
tr_render of the same flattened scene bit for bit; Scene::set_camera between
two renders of one renderer takes effect (the second image equals a fresh
renderer's), a new scene object is uploaded anew, and
create_bdpt_renderer(params)->render (src/factory/renderer.hpp:39-48) equals
tr_render_bdpt bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tracer_amd import abi

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
HARNESS = ROOT / "tests" / "_build" / "facade_harness"
W, H = 64, 48


def _c_abi_image(ctx, seed_scene, camera=None):
    from tracer_amd import native
    s = native.config_scene(1, seed=seed_scene)
    ctx.load_scene(s)
    if camera is not None:
        ctx.set_camera(camera)
    p = s.render_params(width=W, height=H, spp=16, seed=7, filter_type=abi.TR_FILTER_GAUSSIAN, filter_radius=0.5,
                        filter_alpha=0.6)
    _, img, _ = ctx.render(p)
    return img


def test_facade_render_equals_c_abi(gpu_ctx, tmp_path):
    if not HARNESS.exists():
        raise RuntimeError(f"{HARNESS} is not built (run python -m tracer_amd.build)")
    prefix = tmp_path / "img"
    r = subprocess.run([str(HARNESS), str(prefix)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    imgs = {k: np.fromfile(f"{prefix}_{k}.f32", np.float32).reshape(H, W, 3) for k in "abcde"}
    assert all(np.isfinite(v).all() and v.mean() > 0 for v in imgs.values())
    assert np.array_equal(imgs["a"], _c_abi_image(gpu_ctx, 0))
    # set_camera between two renders of the same renderer
    assert not np.array_equal(imgs["a"], imgs["b"])
    assert np.array_equal(imgs["b"], imgs["c"])
    cam = abi.TrCamera()
    cam.pos[:], cam.target[:], cam.up[:] = (0.3, 0.2, 3.0), (0.0, -0.2, 0.0), (0.0, 1.0, 0.0)
    cam.fov_y, cam.lens_radius, cam.focal_distance, cam.aspect = 0.8, 0.0, 1.0, W / H
    assert np.array_equal(imgs["b"], _c_abi_image(gpu_ctx, 0, cam))
    # another scene object on the first renderer
    assert np.array_equal(imgs["d"], _c_abi_image(gpu_ctx, 5))
    # create_bdpt_renderer(...)->render equals tr_render_bdpt of the same scene and camera
    from tracer_amd import native
    s = native.config_scene(1, seed=0)
    gpu_ctx.load_scene(s)
    gpu_ctx.set_camera(cam)
    p = s.bdpt_params(5, 5, width=W, height=H, spp=16, seed=7, filter_type=abi.TR_FILTER_GAUSSIAN, filter_radius=0.5,
                      filter_alpha=0.6)
    _, bimg, _ = gpu_ctx.render_bdpt(p)
    assert np.array_equal(imgs["e"], bimg)


def test_set_camera_rejects_degenerate_cameras(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    for pos, target, up, fov in (((0, 0, 1), (0, 0, 1), (0, 1, 0), 0.7), ((0, 0, 1), (0, 0, 0), (0, 0, 1), 0.7),
                                 ((0, 0, 1), (0, 0, 0), (0, 1, 0), 0.0), ((0, 0, np.nan), (0, 0, 0), (0, 1, 0), 0.7)):
        cam = abi.TrCamera()
        cam.pos[:], cam.target[:], cam.up[:] = pos, target, up
        cam.fov_y, cam.focal_distance, cam.aspect = fov, 1.0, 1.0
        with pytest.raises(native.TracerError):
            gpu_ctx.set_camera(cam)
