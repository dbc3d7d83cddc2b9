"""Image environment light on the GPU (IBL over an HDR texture,
src/light/environment_light.cpp:20-196) against the oracle's restatement
(oracle/refcpu/ref_env.hpp): the same uniforms give the same directions,
pdfs and radiance; lookups agree; and a scene lit by the image renders the
oracle's image in expectation.

Tolerances: directions within 1e-5 (99th percentile); pdf and radiance
within 1e-4 relative away from the poles (|z| < 0.999, where acos of a
float cosine loses the angle's digits) and off cell borders (99.9% of
directions), both sides computing in FP32."""
import numpy as np
import pytest

from tests.test_oracle_env import _blob_image

pytestmark = pytest.mark.gpu


def _env_scene(img, rot=None, detail=1 / 64):
    from tracer_amd import native
    s = native.config_scene(2, detail=detail)
    s.env = (1.0, 1.0, 1.0)
    s.env_image = img
    s.env_world_to_light = rot
    return s


def _rot(a_deg):
    a = np.radians(a_deg)
    return np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])


@pytest.mark.parametrize("rot", [None, 30.0])
def test_env_sampling_matches_oracle(gpu_ctx, rot):
    from oracle import refcpu
    img = _blob_image()
    R = None if rot is None else _rot(rot)
    gpu_ctx.load_scene(_env_scene(img, R))
    ref = refcpu.RefEnv(img, R)
    rng = np.random.default_rng(7)
    u = rng.uniform(size=(1 << 16, 2)).astype(np.float32)
    wi, pdf, L = gpu_ctx.env_sample(u)
    wr, pr, Lr = ref.sample(u)
    assert np.array_equal(pdf > 0, pr > 0)
    ok = pr > 0
    d = np.linalg.norm(wi[ok] - wr[ok], axis=1)
    assert np.quantile(d, 0.99) < 1e-5
    away = ok & (np.abs(wr[:, 2]) < 0.999)
    rp = np.abs(pdf[away] / pr[away] - 1)
    rl = np.abs(L[away] - Lr[away]) / np.maximum(np.abs(Lr[away]), 1e-3)
    print(f"pdf rel: p99.9 {np.quantile(rp, 0.999):.3g}; L rel: p99.9 {np.quantile(rl, 0.999):.3g}")
    assert np.quantile(rp, 0.999) < 1e-4
    assert np.quantile(rl, 0.999) < 1e-4


def test_env_lookup_matches_oracle(gpu_ctx):
    from oracle import refcpu
    img = _blob_image()
    R = _rot(-50.0)
    gpu_ctx.load_scene(_env_scene(img, R))
    ref = refcpu.RefEnv(img, R)
    rng = np.random.default_rng(8)
    d = rng.normal(size=(1 << 16, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    L, pdf = gpu_ctx.env_eval(d)
    Lr, pr = ref.eval(d)
    away = np.abs((d.astype(np.float64) @ R.T)[:, 2]) < 0.999
    rl = np.abs(L[away] - Lr[away]) / np.maximum(np.abs(Lr[away]), 1e-3)
    close = np.isclose(pdf[away], pr[away], rtol=1e-4, atol=0)
    print(f"L rel p99.9 {np.quantile(rl, 0.999):.3g}; pdf agreement {close.mean():.5f}")
    assert np.quantile(rl, 0.999) < 1e-4
    assert close.mean() > 0.999


def test_env_image_scene_matches_fixed_estimator(gpu_ctx):
    """Configuration 2 (one displaced sphere, its lights) under an image
    environment with a bright spot, against the oracle in fixed mode with
    all-lights NEE."""
    from tests.test_gpu_render import _compare_estimators
    img = 0.25 * _blob_image()
    s = _env_scene(img, _rot(20.0))
    s.camera.pos[:] = (0.0, 2.0, 14.0)
    _compare_estimators(gpu_ctx, s, 32, 32, 1024, 128, 8, 8, quirks=False, nee=0, max_depth=6)
