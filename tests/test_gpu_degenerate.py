"""Degenerate inputs through every renderer mode: no NaN or infinity reaches
the film.

The reference turns a non-finite sample into black
(src/renderer/pixel_sampler_renderer.cpp:83-87) and so does the film here
(counted in tr_render_stats::nonfinite_dropped), but a sample that is
non-finite has already lost its light.  These scenes push the estimators to
their singular cases and require that none occurs:

* geometry: zero-area and duplicate triangles, an emitter of zero area (zero
  power, never sampled; a light-sample density of 1 / 0 if it were),
  zero-length and reversed vertex normals;
* materials at their extremes, each on its own block in a lit box: Disney
  with roughness 0 and anisotropy 1, metallic 1, clearcoat gloss 1, full
  transmission with transmission roughness 0 and with ior 1 (a refracted
  direction opposite to wo, whose half vector has length 0), black albedo;
  Phong with exponent 0 and 10^4; glass with ior 1 and with ior 3; Lambert
  with albedo 0 and 1;
* every mode: the plain path tracer (one and three light samples per
  vertex), learned light sampling (guide = 2), the guide network (guide = 1),
  and bidirectional path tracing.
"""
import numpy as np
import pytest

from tracer_amd import abi

pytestmark = pytest.mark.gpu


def _box_with_blocks(materials):
    """A closed room [-2, 2]^3 (Lambert 0.6, inward normals) lit by a ceiling
    quad, with one small block per material along the x axis."""
    from tracer_amd.native import Scene
    P, I, M, E = [], [], [], []

    def box(lo, hi, mat, inward, emit=None):
        c = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])], np.float32)
        quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
        base = sum(len(p) for p in P)
        centre = c.mean(0)
        P.append(c)
        for a, b, cc, d in quads:
            for t in ((a, b, cc), (a, cc, d)):
                p = c[list(t)]
                n = np.cross(p[1] - p[0], p[2] - p[0])
                out = np.dot(n, p.mean(0) - centre) > 0
                I.append(tuple(base + k for k in (t if out != inward else (t[0], t[2], t[1]))))
                M.append(mat)
                E.append((0.0, 0.0, 0.0) if emit is None else emit)

    box((-2, -2, -2), (2, 2, 2), 0, inward=True)
    box((-0.5, 1.9, -0.5), (0.5, 1.95, 0.5), 0, inward=False, emit=(8.0, 8.0, 8.0))
    n = len(materials)
    for k in range(n):
        x0 = -1.8 + 3.6 * k / n
        box((x0 + 0.05, -2.0, -0.6), (x0 + 3.6 / n - 0.05, -0.8, 0.2), 1 + k, inward=False)
    cam = abi.TrCamera()
    cam.pos[:] = (0.0, 0.0, 1.9)
    cam.target[:] = (0.0, -0.8, -1.0)
    cam.up[:] = (0.0, 1.0, 0.0)
    cam.fov_y, cam.focal_distance, cam.aspect = 1.3, 1.0, 1.0
    mats = [abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(0.6, 0.6, 0.6))] + materials
    return Scene(positions=np.concatenate(P), indices=np.array(I, np.int32), material_ids=np.array(M, np.int32),
                 materials=mats, emission=np.array(E, np.float32), camera=cam)


EXTREME = [
    abi.make_material(type=abi.TR_MAT_DISNEY, base_color=(0.8, 0.5, 0.3), roughness=0.0, anisotropic=1.0),
    abi.make_material(type=abi.TR_MAT_DISNEY, base_color=(0.9, 0.9, 0.9), metallic=1.0, roughness=0.0),
    abi.make_material(type=abi.TR_MAT_DISNEY, clearcoat=1.0, clearcoat_gloss=1.0, roughness=0.0, sheen=1.0),
    abi.make_material(type=abi.TR_MAT_DISNEY, transmission=1.0, transmission_roughness=0.0, ior=1.5, roughness=0.0),
    abi.make_material(type=abi.TR_MAT_DISNEY, transmission=1.0, transmission_roughness=0.3, ior=1.0),
    abi.make_material(type=abi.TR_MAT_DISNEY, base_color=(0.0, 0.0, 0.0), roughness=1.0),
    abi.make_material(type=abi.TR_MAT_PHONG, base_color=(0.4, 0.4, 0.4), specular_scale=(0.5, 0.5, 0.5),
                      shininess=0.0),
    abi.make_material(type=abi.TR_MAT_PHONG, base_color=(0.2, 0.2, 0.2), specular_scale=(0.7, 0.7, 0.7),
                      shininess=1e4),
    abi.make_material(type=abi.TR_MAT_GLASS, base_color=(1.0, 1.0, 1.0), specular_scale=(1.0, 1.0, 1.0), ior=1.0),
    abi.make_material(type=abi.TR_MAT_GLASS, base_color=(0.9, 0.9, 0.9), specular_scale=(0.9, 0.9, 0.9), ior=3.0),
    abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(0.0, 0.0, 0.0)),
    abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(1.0, 1.0, 1.0)),
]


def _degenerate_geometry():
    """Zero-area and duplicate triangles, a zero-area emitter, zero-length and
    reversed vertex normals, inside the lit room."""
    s = _box_with_blocks([abi.make_material(type=abi.TR_MAT_DISNEY, roughness=0.2)])
    base = s.positions.shape[0]
    extra = np.array([[0, -1, 0], [1, -1, 0], [0, 0, 0], [2, -1, 0], [3, -1, 0], [0.5, 0.5, 0.5],
                      [-1, -1.5, -1], [-0.2, -1.5, -1], [-1, -0.7, -1]], np.float32)
    tris = np.array([[0, 1, 2], [1, 3, 4], [5, 5, 5], [6, 7, 8], [6, 7, 8], [0, 1, 2]], np.int32) + base
    s.positions = np.concatenate([s.positions, extra])
    s.indices = np.concatenate([s.indices, tris])
    s.material_ids = np.concatenate([s.material_ids, np.ones(len(tris), np.int32)])
    s.emission = np.concatenate([s.emission, np.array([[0, 0, 0], [0, 0, 0], [5, 5, 5], [0, 0, 0], [0, 0, 0],
                                                       [0, 0, 0]], np.float32)])
    nv = s.positions.shape[0]
    N = np.zeros((nv, 3), np.float32)
    N[:base] = (0.0, 1.0, 0.0)
    N[base + 0] = 0.0                       # zero length
    N[base + 1] = (0.0, 0.0, -1.0)          # reversed against the geometric normal
    N[base + 2] = (0.0, 0.0, 1.0)
    N[base + 3:] = (0.0, 0.0, 1.0)
    s.normals = N
    return s


def _modes(ctx, s, W=48, H=48):
    """(name, sums, stats) of every renderer mode on scene s."""
    out = []
    p = dict(width=W, height=H, seed=5, max_depth=8, min_depth=2)
    sums, _, st = ctx.render(s.render_params(spp=32, **p), want_image=False)
    out.append(("plain", sums, st))
    sums, _, st = ctx.render(s.render_params(spp=32, direct_light_sample_num=3, **p), want_image=False)
    out.append(("plain, 3 light samples", sums, st))
    sums, _, st = ctx.render(s.render_params(spp=32, guide=2, guide_light_grid=8, guide_train_passes=2,
                                             samples_per_pass=4, **p), want_image=False)
    out.append(("learned light sampling", sums, st))
    ctx.guide_init(hidden=64, seed=3, batch=1 << 12, steps_per_pass=4)
    sums, _, st = ctx.render(s.render_params(spp=32, guide=1, guide_train_passes=2, samples_per_pass=4, **p),
                             want_image=False)
    out.append(("guide network", sums, st))
    sums, _, st = ctx.render_bdpt(s.bdpt_params(6, 6, width=W, height=H, spp=16, seed=5), want_image=False)
    out.append(("bdpt", sums, st))
    return out


@pytest.mark.parametrize("which", ["materials", "geometry"])
def test_no_nonfinite_samples(gpu_ctx, which):
    s = _box_with_blocks(EXTREME) if which == "materials" else _degenerate_geometry()
    gpu_ctx.load_scene(s)
    bad = []
    for name, sums, st in _modes(gpu_ctx, s):
        finite = bool(np.isfinite(sums).all())
        mean = float(sums[..., :3].sum() / max(sums[..., 3].sum(), 1e-12))
        print(f"{which} / {name}: non-finite samples {st.nonfinite_dropped}, film finite {finite}, mean {mean:.4f}")
        if not finite or st.nonfinite_dropped != 0 or not mean > 0.0:
            bad.append(name)
    assert not bad, f"non-finite samples or film in {bad}"
