"""Bidirectional path tracing (tr_render_bdpt; BDPTRenderer::render,
src/renderer/bdpt_renderer.cpp:1087-1185).

* Closed forms: in the emitting furnace box (tests/furnace.py) a path with k
  scattering vertices carries Le rho^k, so with nc camera and nl light
  vertices the MIS-combined estimator returns Le sum_{k=0}^{nc+nl-2} rho^k,
  and each family of strategies alone (weight 1) returns its own part of the
  sum: s = 0 (camera subpaths that reach a light) k in [0, nc - 2]; s = 1
  (one light vertex) k in [1, nc - 1]; s = S >= 2, k in [S - 1, S + nc - 2];
  t = 1 (light tracing: splats through the camera's importance) k in
  [1, nl - 1]; t = T >= 2, k in [T - 2, T + nl - 2].  A missing or extra
  balance-heuristic term (the reference's light loop skips the strategy with
  no light vertex), a wrong camera importance, density conversion or splat
  weight moves one of these sums.
* Images: configuration 1 against the oracle's path tracer (fixed mode) and
  the OBJ room (Phong block, glass block, fog) against the GPU path tracer,
  each with the maximum path length nc + nl - 2 scattering vertices of the
  bidirectional estimator (max_depth = nc + nl - 2): tile and image z-scores
  over independent seeds.
* Reproducibility (the light-tracing splats are integer sums), progressive
  rendering, the C++ factory entry point and argument errors.
"""
import numpy as np
import pytest

from tests.furnace import furnace_box
from tests.test_gpu_render import _seeds_tiles
from tracer_amd import abi

pytestmark = pytest.mark.gpu

RHO = (0.5, 0.6, 0.7)


def _expected(rho, ks):
    return np.array([sum(r ** k for k in ks) for r in rho])


def _furnace_z(ctx, scene, nc, nl, seeds=8, W=32, H=32, spp=32, **kw):
    ctx.load_scene(scene)
    means = []
    for sd in range(seeds):
        p = scene.bdpt_params(nc, nl, width=W, height=H, spp=spp, seed=100 + sd, **kw)
        sums, img, st = ctx.render_bdpt(p)
        assert st.nonfinite_dropped == 0
        means.append(img.reshape(-1, 3).mean(0))
    means = np.array(means)
    return means.mean(0), means.std(0, ddof=1) / np.sqrt(seeds)


@pytest.mark.parametrize("nc,nl", [(4, 4), (6, 3), (3, 6), (2, 1)])
def test_bdpt_furnace_closed_form(gpu_ctx, nc, nl):
    s = furnace_box(rho=RHO)
    m, se = _furnace_z(gpu_ctx, s, nc, nl)
    want = _expected(RHO, range(nc + nl - 1))
    z = (m - want) / se
    print(f"nc={nc} nl={nl}: mean {m}, expected {want}, z {z}")
    assert np.abs(z).max() < 5.0
    assert np.abs(m / want - 1).max() < 0.01


# (only_s, only_t) -> scattering-vertex counts k of the paths that family samples
def _family_ks(nc, nl, only_s, only_t):
    if only_s:
        S = only_s - 1
        if S == 0:
            return range(0, nc - 1)
        if S == 1:
            return range(1, nc)
        return range(S - 1, S + nc - 1)
    T = only_t
    if T == 1:
        return range(1, nl)
    return range(T - 2, T + nl - 1)


@pytest.mark.parametrize("only_s,only_t", [(1, 0), (2, 0), (3, 0), (4, 0), (0, 1), (0, 2), (0, 3)])
def test_bdpt_strategy_families_closed_form(gpu_ctx, only_s, only_t):
    nc, nl = 4, 4
    s = furnace_box(rho=RHO)
    m, se = _furnace_z(gpu_ctx, s, nc, nl, only_s=only_s, only_t=only_t)
    want = _expected(RHO, _family_ks(nc, nl, only_s, only_t))
    z = (m - want) / se
    rel = np.abs(m / want - 1)
    print(f"only_s={only_s} only_t={only_t}: mean {m}, expected {want}, z {z}, rel {rel}")
    # s = 0 alone has almost no variance here (cosine sampling makes f cos / pdf
    # = rho exactly), so its z-score sees the Lambert sampler's rejection of
    # directions with pdf < 0.001 (diffuse_bxdf.cpp:25-26), a loss of 1e-5
    # per bounce; below 1e-4 relative the mean counts as exact
    assert np.all((np.abs(z) < 5.0) | (rel < 1e-4))
    assert rel.max() < 0.02


def _medium_furnace(sigma_s=0.6, g=0.5):
    """Black walls emitting Le = 1 around a purely scattering medium (albedo 1,
    fog inside every wall): radiance is 1 everywhere, so every pixel is 1 up to
    the paths cut by the vertex budget."""
    s = furnace_box(rho=(0.0, 0.0, 0.0))
    nt = s.indices.shape[0]
    m = abi.TrMedium()
    m.sigma_a[:] = (0.0, 0.0, 0.0)
    m.sigma_s[:] = (sigma_s, sigma_s, sigma_s)
    m.g = g
    m.max_scattering_count = 2 ** 31 - 1
    s.media = [m]
    # every wall faces the interior: its outside (the side of its normal) holds the medium
    s.medium_outside = np.zeros(nt, np.int32)
    s.medium_inside = np.full(nt, -1, np.int32)
    return s


@pytest.mark.parametrize("renderer", ["bdpt", "pt"])
def test_medium_furnace_closed_form(gpu_ctx, renderer):
    s = _medium_furnace()
    if renderer == "bdpt":
        m, se = _furnace_z(gpu_ctx, s, 16, 16)
    else:
        gpu_ctx.load_scene(s)
        vals = []
        for sd in range(8):
            _, img, _ = gpu_ctx.render(s.render_params(width=32, height=32, spp=64, seed=200 + sd, max_depth=30,
                                                       min_depth=64))
            vals.append(img.reshape(-1, 3).mean(0))
        vals = np.array(vals)
        m, se = vals.mean(0), vals.std(0, ddof=1) / np.sqrt(len(vals))
    z = (m - 1.0) / se
    print(f"{renderer}: mean {m}, z {z}")
    assert np.abs(m - 1.0).max() < 0.01 and np.abs(z).max() < 5.0


def _glass_furnace(rho=0.5, lo=(-0.4, -0.4, -0.56), hi=(0.6, 0.5, -0.44)):
    """The emitting furnace box (albedo rho, Le = 1) with a lossless glass slab
    (both colours 1, ior 1.5) between the camera and the far wall.  A uniform
    radiance field passes a lossless dielectric unchanged, so every pixel is
    1 / (1 - rho), up to the paths cut by the vertex budget."""
    from tracer_amd.native import Scene
    f = furnace_box(rho=(rho, rho, rho))
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    centre = corners.mean(0)
    tris = []
    for a, b, c, d in quads:
        for t in ((a, b, c), (a, c, d)):
            p = corners[list(t)]
            n = np.cross(p[1] - p[0], p[2] - p[0])
            tris.append(t if np.dot(n, p.mean(0) - centre) > 0 else (t[0], t[2], t[1]))   # normals outwards
    base = f.positions.shape[0]
    P = np.concatenate([f.positions, corners])
    I = np.concatenate([f.indices, np.array(tris, np.int32) + base])
    nt0, nt1 = f.indices.shape[0], len(tris)
    glass = abi.make_material(type=abi.TR_MAT_GLASS, base_color=(1.0, 1.0, 1.0), specular_scale=(1.0, 1.0, 1.0), ior=1.5)
    mats = list(f.materials) + [glass]
    mid = np.concatenate([f.material_ids, np.ones(nt1, np.int32)])
    em = np.concatenate([f.emission, np.zeros((nt1, 3), np.float32)])
    return Scene(positions=P, indices=I, material_ids=mid, materials=mats, emission=em, camera=f.camera)


@pytest.mark.parametrize("renderer", ["bdpt", "pt"])
def test_glass_furnace_closed_form(gpu_ctx, renderer):
    s = _glass_furnace()
    if renderer == "bdpt":
        m, se = _furnace_z(gpu_ctx, s, 32, 32)
    else:
        gpu_ctx.load_scene(s)
        vals = []
        for sd in range(8):
            _, img, _ = gpu_ctx.render(s.render_params(width=32, height=32, spp=64, seed=300 + sd, max_depth=30,
                                                       min_depth=64))
            vals.append(img.reshape(-1, 3).mean(0))
        vals = np.array(vals)
        m, se = vals.mean(0), vals.std(0, ddof=1) / np.sqrt(len(vals))
    want = 2.0 * (1.0 - 0.5 ** 31)
    z = (m - want) / se
    print(f"{renderer}: mean {m}, expected {want:.6f}, z {z}")
    assert np.abs(m / want - 1.0).max() < 0.005 and np.abs(z).max() < 5.0


def _compare_bdpt(ctx, scene, W, H, bd_spp, seeds, tile, nc, nl, other, z_max=5.0, rel_max=0.02):
    ctx.load_scene(scene)

    def bdpt(seed):
        return ctx.render_bdpt(scene.bdpt_params(nc, nl, width=W, height=H, spp=bd_spp, seed=seed),
                               want_image=False)[0]

    g = _seeds_tiles(bdpt, range(seeds), tile)
    c = _seeds_tiles(other, range(seeds), tile)
    mg, mc = g.mean(0), c.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / seeds + c.var(0, ddof=1) / seeds)
    z = (mg - mc) / np.maximum(se, 1e-6)
    ig, ic = g.mean(axis=(1, 2, 3)), c.mean(axis=(1, 2, 3))
    zi = (ig.mean() - ic.mean()) / np.sqrt(ig.var(ddof=1) / seeds + ic.var(ddof=1) / seeds)
    rel = abs(ig.mean() - ic.mean()) / ic.mean()
    print(f"max tile |z| {np.abs(z).max():.2f}, image z {zi:.2f}, rel {rel:.4f} "
          f"(bdpt {ig.mean():.4f}, path tracer {ic.mean():.4f})")
    assert np.abs(z).max() < z_max and abs(zi) < z_max and rel < rel_max


def test_bdpt_c1_matches_oracle_path_tracer(gpu_ctx):
    from oracle import refcpu
    from tracer_amd import native
    s = native.config_scene(1)
    ref = refcpu.RefScene(s, 3)
    nc = nl = 5
    W = H = 32

    def oracle(seed):
        p = s.render_params(width=W, height=H, spp=128, seed=7000 + seed, max_depth=nc + nl - 2)
        return ref.render(p, quirks=False, nee=0, threads=16)[0]

    _compare_bdpt(gpu_ctx, s, W, H, 256, 8, 8, nc, nl, oracle)


@pytest.mark.parametrize("fog,glass", [(False, False), (True, False), (False, True)])
def test_bdpt_obj_room_matches_path_tracer(gpu_ctx, tmp_path, fog, glass):
    """The OBJ room seen from inside; with fog it is closed on all six sides.
    (With a side open, "fog outside every surface" with the medium of a
    segment decided at its end fills a point with fog only for rays that go
    on to hit a wall; camera and light subpaths then reach different sets of
    paths and the bidirectional weights cannot add up.)  The path tracer does
    not count specular bounces towards its depth, so with glass both path
    lengths are long enough for the cut to be negligible."""
    from tests import objfiles
    from tracer_amd import native
    obj, mtl = objfiles.box_room_obj(glass_block=glass, closed=fog)
    path = objfiles.write(tmp_path, "room.obj", obj, mtl)
    s = native.obj_scene(path, objfiles.camera((0.0, 0.1, 0.95), (0.0, -0.1, -1.0), fov_deg=75.0), outside_fog=fog)
    nc = nl = 32 if glass else 5
    W = H = 32

    def pt(seed):
        p = s.render_params(width=W, height=H, spp=1024, seed=9000 + seed, max_depth=nc + nl - 2, min_depth=64)
        return gpu_ctx.render(p, want_image=False)[0]

    _compare_bdpt(gpu_ctx, s, W, H, 256, 8, 8, nc, nl, pt)


def _env_ground(half=4.0, rho=0.5, env=1.0):
    """A Lambert quad (rho) at y = 0 under a constant environment (env),
    seen from straight above by a narrow pinhole camera: every pixel sees the
    quad, and the quad sees the whole upper hemisphere of sky."""
    from tracer_amd import native
    P = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    cam = abi.TrCamera()
    cam.pos[:] = (0.0, 2.0, 0.0)
    cam.target[:] = (0.0, 0.0, 0.001)
    cam.up[:] = (0.0, 0.0, 1.0)
    cam.fov_y, cam.focal_distance, cam.aspect = 0.7, 1.0, 1.0
    return native.Scene(positions=P, indices=np.array([[0, 2, 1], [0, 3, 2]], np.int32),
                        material_ids=np.zeros(2, np.int32),
                        materials=[abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(rho, rho, rho))],
                        env=(env, env, env), camera=cam)


@pytest.mark.parametrize("only_s", [0, 1, 2, 3])
def test_bdpt_environment_ground_closed_form(gpu_ctx, only_s):
    """Radiance rho L = 0.5 at every pixel, through the weighted
    combination (only_s = 0) and through each strategy alone: camera
    subpaths that escape to the environment (s = 0), light samples of the
    environment (s = 1) and light subpaths that start on the environment's
    disk and are traced to the camera (s = 2).  The last needs the disk on the
    side the light comes from (the reference's IBL::sample_le sends its
    subpaths away from the scene)."""
    s = _env_ground()
    m, se = _furnace_z(gpu_ctx, s, 3, 3, seeds=6, spp=64 if only_s != 3 else 256, only_s=only_s)
    z = (m - 0.5) / np.maximum(se, 1e-9)
    print(f"only_s={only_s}: mean {m}, se {se}, z {z}")
    assert np.abs(z).max() < 5.0 and np.abs(m / 0.5 - 1).max() < 0.02


def test_bdpt_textured_disney_matches_path_tracer(gpu_ctx):
    """Per-hit Disney parameters on both subpaths, over the interpolated
    normals of a UV sphere under an area light and the environment: the
    textured sphere scene against the GPU path tracer."""
    from tests.uvsphere import textured_disney_sphere_scene
    s = textured_disney_sphere_scene(normal_map=False)
    nc = nl = 5
    W = H = 32

    def pt(seed):
        p = s.render_params(width=W, height=H, spp=1024, seed=9000 + seed, max_depth=nc + nl - 2, min_depth=64)
        return gpu_ctx.render(p, want_image=False)[0]

    _compare_bdpt(gpu_ctx, s, W, H, 256, 8, 8, nc, nl, pt)


def test_bdpt_is_reproducible_and_progressive(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    p = s.bdpt_params(5, 5, width=48, height=32, spp=8, seed=3, filter_type=abi.TR_FILTER_BOX, filter_radius=0.5,
                      samples_per_pass=2)
    a, _, _ = gpu_ctx.render_bdpt(p)
    b, _, _ = gpu_ctx.render_bdpt(p)
    assert np.array_equal(a, b)
    one = s.bdpt_params(5, 5, width=48, height=32, spp=8, seed=3, filter_type=abi.TR_FILTER_BOX, filter_radius=0.5,
                        samples_per_pass=8)
    c, _, _ = gpu_ctx.render_bdpt(one)
    np.testing.assert_allclose(c, a, rtol=1e-12, atol=1e-12)
    # two calls of 4 samples accumulate to the same film
    h1 = s.bdpt_params(5, 5, width=48, height=32, spp=4, seed=3, filter_type=abi.TR_FILTER_BOX, filter_radius=0.5)
    f, _, _ = gpu_ctx.render_bdpt(h1)
    h2 = s.bdpt_params(5, 5, width=48, height=32, spp=4, seed=3, filter_type=abi.TR_FILTER_BOX, filter_radius=0.5,
                       sample_offset=4)
    gpu_ctx.render_bdpt(h2, film=f)
    np.testing.assert_allclose(f, a, rtol=1e-9, atol=1e-12)


def test_bdpt_gaussian_filter_matches_box_in_expectation(gpu_ctx):
    """The camera-side estimate goes through the film filter, the splats do
    not: on the uniform furnace image both filters resolve to the same value."""
    s = furnace_box(rho=RHO)
    mg, sg = _furnace_z(gpu_ctx, s, 4, 4, filter_type=abi.TR_FILTER_GAUSSIAN, filter_radius=1.5, filter_alpha=0.6)
    want = _expected(RHO, range(7))
    z = (mg - want) / sg
    print(f"Gaussian r=1.5: mean {mg}, expected {want}, z {z}")
    assert np.abs(z).max() < 5.0


def test_bdpt_arguments(gpu_ctx):
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    for kw in ({"max_camera_vertex_count": 0}, {"max_light_vertex_count": -1}, {"max_camera_vertex_count": 33},
               {"only_s": 12}, {"only_t": 11}, {"filter_radius": 0.0}, {"spp": 0}):
        p = s.bdpt_params(10, 10, width=8, height=8, spp=1, seed=1)
        for k, v in kw.items():
            setattr(p, k, v)
        with pytest.raises(native.TracerError):
            gpu_ctx.render_bdpt(p, want_image=False)
    sun = native.config_scene(5, detail=1 / 2048)
    gpu_ctx.load_scene(sun)
    with pytest.raises(native.TracerError, match="sun"):
        gpu_ctx.render_bdpt(sun.bdpt_params(4, 4, width=8, height=8, spp=1, seed=1), want_image=False)
