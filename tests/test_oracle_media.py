"""Pins of the oracle's heterogeneous medium (oracle/refcpu/ref_pt.cpp
RefGrid, after src/medium/heterogeneous.cpp:10-108 and
LinearSampler::Sample3D, src/utility/sampler.hpp:47-67):

* trilinear interpolation reproduces an affine density field exactly, and
  the lookup repeats with the period of the grid (wrap_repeat);
* ratio tracking has mean exp(-integral of the density) along the segment,
  and delta tracking scatters with probability 1 - that, for a constant grid
  and for a linear gradient (closed-form optical depth);
* the scatter weight is the albedo, or 1 after the first scatter of a run
  with white_for_indirect;
* the facade's create_heterogeneous_medium wiring hands the grid and its
  transform through unchanged.
"""
import numpy as np
import pytest

from tracer_amd import abi

# unit grid cube -> the room [-1, 1]^3
ROOM = np.array([[2.0, 0, 0, -1.0], [0, 2.0, 0, -1.0], [0, 0, 2.0, -1.0]])


def _affine_grid(shape, c0, c):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) / (nz - 1), np.arange(ny) / (ny - 1), np.arange(nx) / (nx - 1), indexing="ij")
    return (c0 + c[0] * x + c[1] * y + c[2] * z).astype(np.float32)


def test_trilinear_reproduces_affine_field_and_wraps():
    from oracle import refcpu
    c0, c = 0.3, (0.5, 0.25, 0.9)
    m = abi.make_grid_medium(_affine_grid((5, 7, 9), c0, c), local_to_world=ROOM)
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.999, 0.999, (2000, 3)).astype(np.float32)
    local = (p + 1) / 2
    want = c0 + local @ np.array(c)
    got = refcpu.grid_density(m, p)
    assert np.max(np.abs(got - want)) < 2e-5
    # one period of the grid in the world is 2 along each axis
    got2 = refcpu.grid_density(m, p + np.float32([2.0, -4.0, 2.0]))
    assert np.max(np.abs(got2 - got)) < 1e-4


@pytest.mark.parametrize("field", ["constant", "gradient"])
def test_tracking_matches_optical_depth(field):
    from oracle import refcpu
    if field == "constant":
        grid = np.full((3, 3, 3), 0.8, np.float32)
    else:
        grid = _affine_grid((4, 4, 6), 0.1, (1.5, 0.0, 0.0))   # density 0.1 + 1.5 x_local
    m = abi.make_grid_medium(grid, albedo=(0.6, 0.5, 0.4), local_to_world=ROOM)
    a, b = np.float32([-0.8, 0.1, 0.2]), np.float32([0.9, 0.1, 0.2])
    L = 1.7
    if field == "constant":
        tau = 0.8 * L
    else:   # x_local runs from 0.1 to 0.95 over the segment, dl = 2 dx_local
        x0, x1 = 0.1, 0.95
        tau = 2 * (0.1 * (x1 - x0) + 0.75 * (x1 ** 2 - x0 ** 2))
    n = 200_000
    tr, sc, w = refcpu.grid_stats(m, a, b, n, seed=3)
    T = np.exp(-tau)
    print(f"{field}: tau {tau:.4f}; tr {tr:.5f} vs {T:.5f}; scatter {sc:.5f} vs {1 - T:.5f}")
    assert abs(sc - (1 - T)) < 4 * np.sqrt(T * (1 - T) / n)
    # the ratio-tracking estimate lies in [0, 1]: its standard deviation is at most 1/2
    assert abs(tr - T) < 4 * 0.5 / np.sqrt(n)
    assert w == pytest.approx(0.6, abs=1e-6)


def test_white_for_indirect_weight():
    from oracle import refcpu
    m = abi.make_grid_medium(np.full((2, 2, 2), 2.0, np.float32), albedo=(0.3, 0.3, 0.3), local_to_world=ROOM,
                             white_for_indirect=True)
    a, b = np.float32([0, 0, -0.9]), np.float32([0, 0, 0.9])
    assert refcpu.grid_stats(m, a, b, 2000, indirect=False)[2] == pytest.approx(0.3, abs=1e-6)
    assert refcpu.grid_stats(m, a, b, 2000, indirect=True)[2] == pytest.approx(1.0, abs=1e-6)


def test_facade_grid_scene(tmp_path):
    from oracle import refcpu
    from tests import objfiles
    from tracer_amd import native
    obj, mtl = objfiles.box_room_obj()
    path = objfiles.write(tmp_path, "room.obj", obj, mtl)
    grid = _affine_grid((4, 5, 6), 0.2, (0.4, 0.3, 0.2))
    cam = objfiles.camera((0.0, 0.0, 3.6), (0.0, 0.0, 0.0))
    s = native.obj_grid_scene(path, cam, grid, ROOM, albedo=(0.7, 0.8, 0.9), g=0.3, max_scattering_count=6)
    assert len(s.media) == 1
    m = s.media[0]
    assert (m.nx, m.ny, m.nz) == (6, 5, 4) and m.max_scattering_count == 6
    assert np.array_equal(m._grid[0], grid)   # the density grid, copied out of the handle
    assert np.allclose(np.array(m.local_to_world[:]).reshape(3, 4), ROOM)
    assert np.allclose(list(m.sigma_s), [0.7, 0.8, 0.9]) and m.g == pytest.approx(0.3)
    assert s.medium_outside is not None and set(np.unique(s.medium_outside)) == {0}
    sums, _, _ = refcpu.RefScene(s, 3).render(s.render_params(width=12, height=12, spp=8, seed=2, max_depth=4),
                                              quirks=False, nee=0, threads=4)
    img = sums[..., :3] / sums[..., 3:]
    assert np.isfinite(img).all() and img.mean() > 0
    with pytest.raises(ValueError):
        native.obj_grid_scene(path, cam, grid[0], ROOM)
    with pytest.raises(native.TracerError):
        native.obj_grid_scene(path, cam, grid, np.zeros((3, 4)))


def test_albedo_and_g_grids():
    """HeterogeneousMedium::sample reads albedo->evaluate and g->evaluate_s at
    the collision (heterogeneous.cpp:66-78): the oracle's lookups of albedo
    and g grids reproduce affine fields exactly (trilinear interpolation of an
    affine function), on grids of their own sizes; constant grids give the
    constant's scatter weight."""
    from oracle import refcpu
    rng = np.random.default_rng(4)
    dens = rng.uniform(0.5, 1.5, (3, 4, 5)).astype(np.float32)
    alb = np.stack([_affine_grid((6, 5, 4), 0.2, (0.3, 0.1, 0.2)), _affine_grid((6, 5, 4), 0.5, (-0.2, 0.2, 0.1)),
                    _affine_grid((6, 5, 4), 0.4, (0.1, -0.1, 0.3))], -1)
    g = _affine_grid((4, 7, 3), -0.3, (0.4, 0.2, -0.1))
    m = abi.make_grid_medium(dens, albedo=alb, g=g, local_to_world=ROOM)
    p = rng.uniform(-0.999, 0.999, (2000, 3)).astype(np.float32)
    local = (p + 1) / 2
    a, gg = refcpu.grid_albedo_g(m, p)
    want_a = np.stack([0.2 + local @ np.array([0.3, 0.1, 0.2]), 0.5 + local @ np.array([-0.2, 0.2, 0.1]),
                       0.4 + local @ np.array([0.1, -0.1, 0.3])], -1)
    assert np.max(np.abs(a - want_a)) < 2e-5
    assert np.max(np.abs(gg - (-0.3 + local @ np.array([0.4, 0.2, -0.1])))) < 2e-5
    # without grids: the constants
    m2 = abi.make_grid_medium(dens, albedo=(0.7, 0.6, 0.5), g=0.25, local_to_world=ROOM)
    a2, g2 = refcpu.grid_albedo_g(m2, p)
    assert np.all(a2 == np.float32([0.7, 0.6, 0.5])) and np.all(g2 == np.float32(0.25))
    # a constant albedo grid scatters with the constant's weight
    a_const = np.full((2, 3, 2, 3), 0.6, np.float32)
    m3 = abi.make_grid_medium(dens, albedo=a_const, local_to_world=ROOM)
    pa, pb = np.float32([-0.9, 0.1, 0.2]), np.float32([0.9, -0.2, -0.1])
    _, frac, w = refcpu.grid_stats(m3, pa, pb, 20000, seed=3)
    assert frac > 0.5 and abs(w - 0.6) < 1e-6
