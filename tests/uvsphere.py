"""A lit scene with a UV-mapped sphere: vertex normals and texture
coordinates on every vertex, so that the shading frame of a hit follows the
reference's normal derivatives (src/shape/triangle.cpp:200-211)."""
from __future__ import annotations

import numpy as np

from tracer_amd import abi


def uv_sphere(n_phi: int, n_theta: int, centre=(0.0, 0.0, 0.0), radius=1.0):
    """Latitude-longitude sphere: vertex (i, j) at phi = 2 pi i / n_phi,
    theta = pi j / n_theta, uv = (i / n_phi, j / n_theta), normal = the unit
    radial direction.  Returns positions, normals, uvs, triangle indices."""
    i, j = np.meshgrid(np.arange(n_phi + 1), np.arange(n_theta + 1), indexing="ij")
    phi = 2 * np.pi * i / n_phi
    th = np.pi * j / n_theta
    nrm = np.stack([np.sin(th) * np.cos(phi), np.cos(th), np.sin(th) * np.sin(phi)], -1).reshape(-1, 3)
    pos = np.asarray(centre) + radius * nrm
    uv = np.stack([i / n_phi, j / n_theta], -1).reshape(-1, 2)
    vid = lambda a, b: a * (n_theta + 1) + b
    tris = []
    for a in range(n_phi):
        for b in range(n_theta):
            p00, p10, p01, p11 = vid(a, b), vid(a + 1, b), vid(a, b + 1), vid(a + 1, b + 1)
            # wound so that the geometric normal points outwards, with the vertex normals
            if b > 0:
                tris.append((p00, p10, p01))
            if b < n_theta - 1:
                tris.append((p10, p11, p01))
    return pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32), np.array(tris, np.int32)


def anisotropic_sphere_scene(material, light=(0.0, 2.5, 0.0), env=0.3):
    """The sphere (material 0) above a grey ground quad (material 1), lit by a
    downward-facing square emitter and a constant environment; camera on +z."""
    from tracer_amd.native import Scene
    P, N, UV, T = uv_sphere(48, 24)
    parts = [(P, N, UV, T, 0, None)]

    def quad(c, ex, ey, mat, emit=None):
        c, ex, ey = map(np.asarray, (c, ex, ey))
        p = np.array([c - ex - ey, c + ex - ey, c + ex + ey, c - ex + ey], np.float32)
        n = np.cross(ex, ey)
        n = np.tile(n / np.linalg.norm(n), (4, 1)).astype(np.float32)
        uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
        return p, n, uv, np.array([[0, 1, 2], [0, 2, 3]], np.int32), mat, emit

    parts.append(quad((0, -1.0, 0), (0, 0, 4), (4, 0, 0), 1))                    # ground, normal +y
    parts.append(quad(light, (0.6, 0, 0), (0, 0, 0.6), 2, (12.0, 12.0, 12.0)))   # emitter, normal -y
    pos, nrm, uvs, idx, mid, emis = [], [], [], [], [], []
    base = 0
    for p, n, uv, t, m, e in parts:
        pos.append(p); nrm.append(n); uvs.append(uv); idx.append(t + base)
        mid.append(np.full(len(t), m, np.int32))
        emis.append(np.tile(np.asarray(e if e is not None else (0.0, 0.0, 0.0), np.float32), (len(t), 1)))
        base += len(p)
    cam = abi.TrCamera()
    cam.pos[:] = (0.0, 0.6, 4.0)
    cam.target[:] = (0.0, -0.2, 0.0)
    cam.up[:] = (0.0, 1.0, 0.0)
    cam.fov_y, cam.focal_distance, cam.aspect = 0.7, 1.0, 1.0
    mats = [material, abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(0.5, 0.5, 0.5)),
            abi.make_material(type=abi.TR_MAT_LAMBERT, base_color=(0.0, 0.0, 0.0))]
    return Scene(positions=np.concatenate(pos), indices=np.concatenate(idx), material_ids=np.concatenate(mid),
                 materials=mats, normals=np.concatenate(nrm), uvs=np.concatenate(uvs), emission=np.concatenate(emis),
                 env=(env, env, env), camera=cam)


def textured_disney_sphere_scene(normal_map=True, seed=0):
    """The sphere scene with a Disney material whose base colour, metallic,
    roughness, specular scale, clearcoat and sheen come from float textures
    (stripes and checks in u and v) and, optionally, a bump normal map
    (tangent-space normals tilted by up to ~30 degrees, encoded as (n + 1) / 2).
    Returns the scene; textures are numbered from 1 in Scene.textures."""
    rng = np.random.default_rng(seed)
    h, w = 32, 64
    v, u = np.meshgrid((np.arange(h) + 0.5) / h, (np.arange(w) + 0.5) / w, indexing="ij")
    base = np.stack([0.2 + 0.7 * u, 0.3 + 0.5 * v, 0.8 - 0.6 * u], -1)
    checks = ((np.floor(u * 8) + np.floor(v * 4)) % 2)
    metallic = np.repeat(checks[..., None], 3, -1)
    rough = np.repeat((0.1 + 0.8 * (0.5 + 0.5 * np.sin(2 * np.pi * 3 * u)))[..., None], 3, -1)
    spec = np.stack([0.5 + 0.5 * v, 0.5 + 0.3 * u, np.full_like(u, 0.8)], -1)
    coat = np.repeat((v > 0.5).astype(float)[..., None], 3, -1)
    sheen = np.repeat((0.5 * u)[..., None], 3, -1)
    texs = [base, metallic, rough, spec, coat, sheen]
    if normal_map:
        tx = 0.6 * np.sin(2 * np.pi * 6 * u)
        ty = 0.6 * np.sin(2 * np.pi * 3 * v)
        n = np.stack([tx, ty, np.ones_like(u)], -1)
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        texs.append((n + 1) / 2)
    mat = abi.make_material(type=abi.TR_MAT_DISNEY, base_color_texture=1, roughness=0.5, transmission=0.0, ior=1.5,
                            sheen_tint=0.5, clearcoat_gloss=0.7,
                            param_textures={"metallic": 2, "roughness": 3, "specular_scale": 4, "clearcoat": 5,
                                            "sheen": 6},
                            normal_texture=7 if normal_map else 0)
    s = anisotropic_sphere_scene(mat)
    s.textures = [t.astype(np.float32) for t in texs]
    return s
