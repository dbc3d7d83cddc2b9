"""Device lookups against the oracle point by point: the density of a grid
medium (trilinear over wrapped grid coordinates, src/utility/sampler.hpp:47-67)
and 8-bit texture lookups (bilinear, src/core/texture.hpp:57-83, fixed mode:
texels interpolated as floats).  The oracle interpolates in double, the
device in float (with fused multiply-adds): agreement within 1e-5 of the
largest value, where a wrong voxel, axis or wrap is off by a whole texel step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOM = np.array([[2.0, 0, 0, -1.0], [0, 2.0, 0, -1.0], [0, 0, 2.0, -1.0]])


def _room(tmp_path, texture=None):
    from tests import objfiles
    obj, mtl = objfiles.box_room_obj(texture=texture)
    return objfiles.write(tmp_path, "room.obj", obj, mtl), objfiles.camera((0.0, 0.0, 3.6), (0.0, 0.0, 0.0))


def test_grid_density_matches_oracle(gpu_ctx, tmp_path):
    from oracle import refcpu
    from tracer_amd import native
    rng = np.random.default_rng(21)
    grid = rng.uniform(0.0, 3.0, (7, 9, 11)).astype(np.float32)
    path, cam = _room(tmp_path)
    rot = np.array([[0.0, -1.6, 0.0, 0.8], [2.0, 0.0, 0.0, -1.0], [0.0, 0.0, 2.5, -1.2]])   # rotated, scaled, shifted
    for l2w in (ROOM, rot):
        s = native.obj_grid_scene(path, cam, grid, l2w, albedo=(0.9, 0.9, 0.9), g=0.2)
        gpu_ctx.load_scene(s)
        pts = rng.uniform(-3.0, 3.0, (1 << 16, 3)).astype(np.float32)   # inside the room and repeated outside
        got = gpu_ctx.medium_density(0, pts)
        want = refcpu.grid_density(s.media[0], pts)
        err = np.abs(got - want).max()
        print(f"max |density difference| {err:.3g}")
        assert err < 1e-5 * grid.max()
    with pytest.raises(native.TracerError):
        gpu_ctx.medium_density(1, pts[:4])


def test_grid_albedo_and_g_match_oracle(gpu_ctx, tmp_path):
    """Device albedo and asymmetry of a grid medium (tr_medium_albedo_g_batch)
    against the oracle's lookups, for albedo and g grids of their own sizes
    under a rotated transform, and for the constants."""
    from oracle import refcpu
    from tracer_amd import abi, native
    rng = np.random.default_rng(23)
    path, cam = _room(tmp_path)
    dens = rng.uniform(0.0, 2.0, (5, 6, 7)).astype(np.float32)
    alb = rng.uniform(0.0, 1.0, (4, 9, 6, 3)).astype(np.float32)
    g = rng.uniform(-0.95, 0.95, (7, 3, 5)).astype(np.float32)
    rot = np.array([[0.0, -1.6, 0.0, 0.8], [2.0, 0.0, 0.0, -1.0], [0.0, 0.0, 2.5, -1.2]])
    s = native.obj_grid_scene(path, cam, dens, rot)
    for albedo, gg in ((alb, g), ((0.3, 0.5, 0.7), 0.35)):
        s.media[0] = abi.make_grid_medium(dens, albedo=albedo, g=gg, local_to_world=rot)
        gpu_ctx.load_scene(s)
        pts = rng.uniform(-3.0, 3.0, (1 << 16, 3)).astype(np.float32)
        ga, g_g = gpu_ctx.medium_albedo_g(0, pts)
        oa, o_g = refcpu.grid_albedo_g(s.media[0], pts)
        err = max(np.abs(ga - oa).max(), np.abs(g_g - np.clip(o_g, -1 + 3e-4, 1 - 3e-4)).max())
        print(f"max |albedo, g difference| {err:.3g}")
        assert err < 1e-5


def test_texture_lookup_matches_oracle(gpu_ctx, tmp_path):
    from oracle import refcpu
    from tracer_amd import native
    rng = np.random.default_rng(22)
    img = rng.integers(0, 256, (13, 17, 3), dtype=np.uint8)
    native.write_png(tmp_path / "t.png", img)
    path, cam = _room(tmp_path, texture="t.png")
    s = native.obj_scene(path, cam, outside_fog=False)
    gpu_ctx.load_scene(s)
    uv = rng.uniform(-2.0, 3.0, (1 << 16, 2)).astype(np.float32)
    for t, tex in enumerate(s.textures, start=1):
        got = gpu_ctx.texture_eval(t, uv)
        want = refcpu.tex_eval8(tex, uv, quirks=False)
        err = np.abs(got - want).max()
        print(f"texture {t}: max |difference| {err:.3g}")
        assert err < 1e-5
    with pytest.raises(native.TracerError):
        gpu_ctx.texture_eval(0, uv[:4])
    with pytest.raises(native.TracerError):
        gpu_ctx.texture_eval(len(s.textures) + 1, uv[:4])
