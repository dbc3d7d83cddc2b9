"""GPU parity of the film, camera, light-selection and phase-function device
code against the oracle, and of the whole estimator against closed forms.

* film: k_accumulate (tr_film_accumulate_batch) against the oracle's
  Tile::add_sample (src/core/render.hpp:37-57) at the sample positions the
  renderer draws (its RNG stream, replayed on the host by tr_rng_sequence),
  Gaussian filters of radius 0.5 to 2.5 and box filters, with non-finite
  samples counted as black (path_trace_renderer.cpp:199-203);
* camera: device rays against the oracle's ThinLensCamera for a pinhole and a
  thin lens;
* shading frames: the device's tangent frame at surface hits against the
  oracle's SurfaceIntersection frame, with and without texture coordinates;
* light selection: the device pmf and pick_light against the oracle's power
  distribution and Distribution1D::sample_discrete;
* Henyey-Greenstein: same scattering cosine and pdf as the oracle for the same
  uniforms, and a chi-square of the device samples;
* furnace: the device renderer's image of a closed emitting Lambertian box is
  Le * sum_{k<=D} rho^k, with and without Russian roulette;
* configuration 1 through the reference's Gaussian filter (0.5, 0.6) in
  reference mode.
"""
import numpy as np
import pytest

from tests.furnace import furnace_box, furnace_expected
from tracer_amd import abi

pytestmark = pytest.mark.gpu


def _sample_positions(seed, W, H, spp, offset):
    from tracer_amd import native
    lib = native.hip_lib()
    import ctypes as C
    lib.tr_rng_sequence.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.POINTER(C.c_uint32)]
    lib.tr_rng_sequence.restype = None
    pos = np.zeros((spp, H, W, 2), np.float32)
    buf = np.zeros(2, np.uint32)
    for s in range(spp):
        for y in range(H):
            for x in range(W):
                lib.tr_rng_sequence(seed, y * W + x, offset + s, 2, buf.ctypes.data_as(C.POINTER(C.c_uint32)))
                u = (buf >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
                pos[s, y, x] = (np.float32(x) + u[0], np.float32(y) + u[1])
    return pos


@pytest.mark.parametrize("ftype,r,alpha", [(abi.TR_FILTER_GAUSSIAN, 0.5, 0.6), (abi.TR_FILTER_GAUSSIAN, 1.0, 0.6),
                                           (abi.TR_FILTER_GAUSSIAN, 1.5, 2.0), (abi.TR_FILTER_GAUSSIAN, 2.5, 0.3),
                                           (abi.TR_FILTER_BOX, 0.5, 0.0), (abi.TR_FILTER_BOX, 1.2, 0.0)])
def test_film_accumulate_matches_oracle_splat(gpu_ctx, ftype, r, alpha):
    from oracle import refcpu
    W, H, spp, seed, offset = 19, 13, 5, 77, 3
    rng = np.random.default_rng(int(r * 10) + ftype)
    L = rng.uniform(0, 4, (spp, H, W, 3)).astype(np.float32)
    L[1, 2, 3] = (np.nan, 1.0, 1.0)
    L[3, 7, 0, 2] = np.inf
    p = abi.TrRenderParams()
    p.width, p.height, p.spp, p.seed, p.sample_offset = W, H, spp, seed, offset
    p.filter_type, p.filter_radius, p.filter_alpha = ftype, r, alpha
    gpu = gpu_ctx.film_accumulate(p, L)
    pos = _sample_positions(seed, W, H, spp, offset)
    Lc = np.where(np.isfinite(L).all(-1, keepdims=True), L, 0.0)   # invalid samples are black, weight kept
    ref = refcpu.film_splat(W, H, ftype, r, alpha, pos.reshape(-1, 2), Lc.reshape(-1, 3))
    assert (ref[..., 3] > 0).all()
    assert np.allclose(gpu, ref, rtol=2e-6, atol=1e-9), np.abs(gpu - ref).max()


def test_render_counts_nonfinite_as_black(gpu_ctx):
    """The image of a scene whose every sample is finite does not change when
    the filter gathers across pixels; sums and weights stay consistent."""
    from tracer_amd import native
    s = native.config_scene(1)
    gpu_ctx.load_scene(s)
    p = s.render_params(width=32, height=32, spp=8, seed=5, filter_type=abi.TR_FILTER_GAUSSIAN, filter_radius=1.5,
                        filter_alpha=0.6)
    sums, img, st = gpu_ctx.render(p)
    assert st.nonfinite_dropped == 0 and np.isfinite(sums).all() and (sums[..., 3] > 0).all()


def test_camera_rays_match_oracle(gpu_ctx):
    from oracle import refcpu
    from tracer_amd import native
    rng = np.random.default_rng(3)
    film = rng.uniform(0, 1, (4096, 2)).astype(np.float32)
    lens = rng.uniform(0, 1, (4096, 2)).astype(np.float32)
    for lens_radius, focal, aspect in ((0.0, 1.0, 1.0), (0.25, 6.0, 1.5)):
        s = native.config_scene(1)
        s.camera.lens_radius = lens_radius
        s.camera.focal_distance = focal
        s.camera.aspect = aspect
        gpu_ctx.load_scene(s)
        g = gpu_ctx.camera_rays(film, lens)
        o = refcpu.RefScene(s, 3).camera_rays(film, lens)
        assert np.allclose(g[:, :3], o[:, :3], atol=2e-6 * max(1.0, np.abs(o[:, :3]).max()))
        assert np.allclose(g[:, 4:7], o[:, 4:7], atol=2e-6)
        assert np.allclose(np.linalg.norm(g[:, 4:7], axis=1), 1.0, atol=1e-6)


@pytest.mark.parametrize("scene", ["c1", "uv_sphere"])
def test_shading_frames_match_oracle(gpu_ctx, scene):
    """Device shading frames (tr_shading_frame_batch) against the oracle's
    SurfaceIntersection frames (triangle.cpp:178-211): coordinate(sn) on
    configuration 1 (no texture coordinates), the normal derivatives of
    compute_ss_ts on a UV-mapped sphere with an anisotropic material."""
    from oracle import refcpu
    from tests.uvsphere import anisotropic_sphere_scene
    from tracer_amd import native
    if scene == "c1":
        s = native.config_scene(1)
    else:
        s = anisotropic_sphere_scene(abi.make_material(type=abi.TR_MAT_DISNEY, metallic=1.0, roughness=0.3,
                                                       anisotropic=0.9))
    gpu_ctx.load_scene(s)
    rng = np.random.default_rng(11)
    n = 8192
    prims = rng.integers(0, s.num_triangles, n).astype(np.int32)
    b = rng.dirichlet((1, 1, 1), n)[:, 1:].astype(np.float32)
    g = gpu_ctx.shading_frames(prims, b)
    o = refcpu.RefScene(s, 3).shading_frames(prims, b)
    err = np.abs(g - o).max(axis=(1, 2))
    print(f"{scene}: max frame difference {err.max():.2e}")
    assert err.max() < 2e-5


@pytest.mark.parametrize("config,detail", [(1, 1.0), (2, 1 / 16), (5, 1 / 2048)])
def test_light_selection_matches_oracle(gpu_ctx, config, detail):
    from oracle import refcpu
    from tracer_amd import native
    s = native.config_scene(config, detail=detail)
    gpu_ctx.load_scene(s)
    ref = refcpu.RefScene(s, 3)
    pg = gpu_ctx.light_pmf()
    po = ref.light_pmf(len(pg) - 2)
    assert np.allclose(pg, po, rtol=2e-6, atol=1e-9)
    u = np.random.default_rng(9).uniform(0, 1, 100000).astype(np.float32)
    ig, mg = gpu_ctx.light_pick(u)
    io, mo = ref.light_pick(u)
    cdf = np.cumsum(po.astype(np.float64))
    keep = np.min(np.abs(u[:, None].astype(np.float64) - cdf[None, :]), axis=1) > 1e-5
    assert np.array_equal(ig[keep], io[keep])
    assert np.allclose(mg, po[ig], rtol=1e-6)


@pytest.mark.parametrize("g", [0.7, -0.4, 0.0, 0.3])
def test_hg_sampler_matches_oracle(gpu_ctx, g):
    from tests.test_oracle_pins import _hg_samples, hg_chi2
    wo, u, wi_o, pdf_o = _hg_samples(g, 200000, 11)
    wi, pdf = gpu_ctx.hg_sample(g, wo, u)
    # the two frames around wo differ in azimuth; the scattering cosine and pdf agree
    c_g = np.einsum("ij,ij->i", wi.astype(np.float64), wo)
    c_o = np.einsum("ij,ij->i", wi_o.astype(np.float64), wo)
    assert np.abs(c_g - c_o).max() < 2e-5
    assert np.allclose(pdf, pdf_o, rtol=1e-4)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=1e-5)
    p_mu, p_phi = hg_chi2(g, wo, wi)
    assert p_mu > 1e-4 and p_phi > 1e-4


@pytest.mark.parametrize("max_depth,min_depth", [(4, 10), (6, 1), (12, 3)])
def test_furnace_closed_form(gpu_ctx, max_depth, min_depth):
    rho = (0.5, 0.6, 0.7)
    s = furnace_box(rho=rho)
    gpu_ctx.load_scene(s)
    p = s.render_params(width=64, height=64, spp=256, seed=8, max_depth=max_depth, min_depth=min_depth)
    sums, _, st = gpu_ctx.render(p, want_image=False)
    img = sums[..., :3] / sums[..., 3:]
    px = img.reshape(-1, 3)
    m, se = px.mean(0), px.std(0, ddof=1) / np.sqrt(px.shape[0])
    z = (m - furnace_expected((1, 1, 1), rho, max_depth)) / se
    print(f"furnace D={max_depth}: mean {m}, expected {furnace_expected((1, 1, 1), rho, max_depth)}, z {z}")
    assert np.abs(z).max() < 5.0
    for wrong in (max_depth - 1, max_depth + 1):
        assert np.abs((m - furnace_expected((1, 1, 1), rho, wrong)) / se).max() > 10.0


def test_c1_gaussian_filter_matches_reference_estimator(gpu_ctx):
    """Configuration 1 through create_gaussin_filter(0.5, 0.6), the reference's
    only filter, against the oracle in reference mode."""
    from tests.test_gpu_render import _compare_estimators
    from tracer_amd import native
    s = native.config_scene(1)
    s.defaults.filter_type = abi.TR_FILTER_GAUSSIAN
    s.defaults.filter_radius = 0.5
    s.defaults.filter_alpha = 0.6
    _compare_estimators(gpu_ctx, s, 32, 32, 256, 32, 8, 8, quirks=True, nee=0, rel_max=0.01)
