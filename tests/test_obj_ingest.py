"""OBJ / MTL ingest through the C++ facade (load_model_from_file,
create_texture_from_file, the run()-style wiring of src/main.cpp:212-282),
checked on the flattened scene it produces.

The reference reads OBJ files with tinyobjloader, which it does not vendor
and whose version it does not pin (dep/tinyobjloader.cmake), and it ships no
OBJ assets.  These tests pin the reader by construction: hand-written files
whose triangles, areas, normals, materials and lights are known, the same
geometry written with every corner syntax, and the error cases."""
import numpy as np
import pytest

from tests import objfiles
from tracer_amd import abi, native

CAM = objfiles.camera((0.0, 0.0, 3.6), (0.0, 0.0, 0.0))


def _tris(s):
    P = s.positions.reshape(-1, 3).astype(np.float64)
    return P[s.indices.reshape(-1, 3)]


def _areas(T):
    return 0.5 * np.linalg.norm(np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0]), axis=1)


MTL = "newmtl a\nKd 0.5 0.4 0.3\nnewmtl b\nKd 0.2\nKs 0.3 0.3 0.3\nNs 40\n"


def _load(tmp_path, obj, mtl=None, fog=False):
    if mtl is None:   # every triangle needs a material (main.cpp:263): a default one
        obj, mtl = "mtllib scene.mtl\nusemtl a\n" + obj, MTL
    return native.obj_scene(objfiles.write(tmp_path, "m.obj", obj, mtl), CAM, outside_fog=fog)


def test_room_scene_geometry_materials_lights(tmp_path):
    obj, mtl = objfiles.box_room_obj()
    s = native.obj_scene(objfiles.write(tmp_path, "room.obj", obj, mtl), CAM, outside_fog=False)
    T = _tris(s)
    # 5 room quads + 1 light quad + 2 blocks x 5 quads = 16 quads = 32 triangles
    assert s.num_triangles == 32
    area = _areas(T).sum()
    tall = 0.6 * 0.6 + 4 * 0.6 * (0.25 + 0.998)   # top and four sides
    short = 0.5 * 0.5 + 4 * 0.5 * (-0.4 + 0.998)
    assert abs(area - (5 * 4.0 + 0.25 + tall + short)) < 1e-4
    # every quad of the room keeps 4 distinct (position, normal) corners: 16 quads x 4
    assert s.positions.size // 3 == 64
    # Phong with ks = 0 -> Lambert with kd, Phong with ks -> TR_MAT_PHONG (one per MTL material in use)
    assert len(s.materials) == 5
    phong = [m for m in s.materials if m.type == abi.TR_MAT_PHONG]
    assert len(phong) == 1 and all(m.type == abi.TR_MAT_LAMBERT for m in s.materials if m.type != abi.TR_MAT_PHONG)
    glossy = phong[0]
    assert np.allclose(glossy.base_color[:], [0.45, 0.40, 0.30]) and np.allclose(glossy.specular_scale[:], 0.35)
    assert glossy.shininess == 25.0
    assert any(np.allclose(m.base_color[:], [0.65, 0.05, 0.05]) for m in s.materials)
    # the tall block's 10 triangles use the glossy material
    gid = next(i for i, m in enumerate(s.materials) if m.type == abi.TR_MAT_PHONG)
    assert (s.material_ids == gid).sum() == 10
    # the two light triangles carry Ke, nothing else emits
    em = s.emission.reshape(-1, 3)
    lit = np.nonzero(em.sum(1) > 0)[0]
    assert len(lit) == 2 and np.allclose(em[lit], 12.0)
    assert np.allclose(_areas(T[lit]).sum(), 0.25)
    # light triangles wind towards -y (one-sided emitters face the room)
    nrm = np.cross(T[lit, 1] - T[lit, 0], T[lit, 2] - T[lit, 0])
    assert np.all(nrm[:, 1] < 0)
    # shading normals are the given vn, normalised
    N = s.normals.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-6)


def test_fog_outside_and_vacuum_inside(tmp_path):
    obj, mtl = objfiles.box_room_obj()
    s = native.obj_scene(objfiles.write(tmp_path, "room.obj", obj, mtl), CAM, outside_fog=True)
    assert len(s.media) == 1
    m = s.media[0]
    assert np.allclose(m.sigma_a[:], 0.03) and np.allclose(m.sigma_s[:], 0.3) and abs(m.g - 0.8) < 1e-7
    assert m.max_scattering_count == 25
    assert np.all(s.medium_outside == 0) and np.all(s.medium_inside == -1)


CORNER_FORMS = {
    "v": "f 1 2 3 4",
    "v/vt": "f 1/1 2/2 3/3 4/4",
    "v//vn": "f 1//1 2//1 3//1 4//1",
    "v/vt/vn": "f 1/1/1 2/2/1 3/3/1 4/4/1",
    "negative": "f -4/-4/-1 -3/-3/-1 -2/-2/-1 -1/-1/-1",
}


@pytest.mark.parametrize("form", list(CORNER_FORMS))
def test_corner_syntaxes_give_the_same_triangles(tmp_path, form):
    head = ("mtllib scene.mtl\nv 0 0 0\nv 2 0 0\nv 2 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
            "vn 0 0 1\nusemtl a\n")
    s = _load(tmp_path, head + CORNER_FORMS[form] + "\n", MTL)
    T = _tris(s)
    assert s.num_triangles == 2 and abs(_areas(T).sum() - 2.0) < 1e-6
    N = s.normals.reshape(-1, 3)
    assert np.allclose(N, [0, 0, 1])   # given or geometric, both +z here
    if "vt" in form or form == "negative":
        uv = s.uvs.reshape(-1, 2)
        P = s.positions.reshape(-1, 3)
        assert np.allclose(uv, P[:, :2] / [2, 1])


def test_quad_splits_along_shorter_diagonal(tmp_path):
    # |p2 - p0| = |(3,1)| < |p3 - p1| = |(-3.5,1)|: triangles (0,1,2), (0,2,3)
    obj = "v 0 0 0\nv 3 0 0\nv 3 1 0\nv -0.5 1 0\nf 1 2 3 4\n"
    s = _load(tmp_path, obj + "", None)
    T = _tris(s)
    diag = {tuple(sorted(map(tuple, t.round(6).tolist()))) for t in T}
    assert ((0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (3.0, 1.0, 0.0)) in diag
    # mirrored: |p2 - p0| > |p3 - p1|: triangles (0,1,3), (1,2,3)
    s = _load(tmp_path, "v 0 0 0\nv 3 0 0\nv 3.5 1 0\nv 0 1 0\nf 1 2 3 4\n")
    T = _tris(s)
    diag = {tuple(sorted(map(tuple, t.round(6).tolist()))) for t in T}
    assert ((0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (3.0, 0.0, 0.0)) in diag


def test_polygon_fan_and_face_normals(tmp_path):
    # a convex hexagon in the plane z = 1 wound counter-clockwise seen from +z, no normals
    ang = np.arange(6) * np.pi / 3
    pts = np.stack([np.cos(ang), np.sin(ang), np.ones(6)], 1)
    obj = "".join(f"v {x} {y} {z}\n" for x, y, z in pts) + "f 1 2 3 4 5 6\n"
    s = _load(tmp_path, obj)
    assert s.num_triangles == 4
    assert abs(_areas(_tris(s)).sum() - 1.5 * np.sqrt(3)) < 1e-5
    assert np.allclose(s.normals.reshape(-1, 3), [0, 0, 1], atol=1e-6)


def test_groups_are_separate_meshes(tmp_path):
    """Vertices are deduplicated within a shape, not across shapes: a corner
    shared by two groups is stored twice."""
    obj = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\no first\nf 1 2 3\nf 2 4 3\n"
    s = _load(tmp_path, obj)
    assert s.positions.size // 3 == 4
    s = _load(tmp_path, obj.replace("f 2 4 3", "g second\nf 2 4 3"))
    assert s.positions.size // 3 == 6


@pytest.mark.parametrize("ks,want", [("", (1.0, 1.0, 1.0)), ("Ks 0.5 0.6 0.7\n", (0.5, 0.6, 0.7))])
def test_ior_above_one_makes_glass(tmp_path, ks, want):
    """image_file.cpp:156-163: Ni > 1 gives a glass with Ks (white when Ks is
    black) as both reflection and transmission colour."""
    obj = "mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl g\nf 1 2 3\n"
    s = _load(tmp_path, obj, "newmtl g\nKd 0.8 0.8 0.8\n" + ks + "Ni 1.5\n")
    g = s.materials[0]
    assert g.type == abi.TR_MAT_GLASS and abs(g.ior - 1.5) < 1e-7
    assert np.allclose(g.base_color[:], want) and np.allclose(g.specular_scale[:], want)


def test_mtl_values(tmp_path):
    obj = "mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl b\nf 1 2 3\n"
    s = _load(tmp_path, obj, MTL)
    assert len(s.materials) == 1   # the flattened scene keeps the materials in use
    b = s.materials[0]
    # "Kd 0.2" sets only the first component
    assert b.type == abi.TR_MAT_PHONG and np.allclose(b.base_color[:], [0.2, 0, 0]) and b.shininess == 40.0
    assert np.allclose(b.specular_scale[:], 0.3) and s.material_ids.tolist() == [0]


@pytest.mark.parametrize("obj,mtl,msg", [
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", None, "invalid face corner"),
    ("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n", None, "invalid face corner"),
    ("v 0 0\n", None, "vertex needs 3 coordinates"),
    ("mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", MTL, "has no material"),
    ("mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl t\nf 1 2 3\n", "newmtl t\nmap_Kd -bm 1 wood.png\n",
     "image wood.png: cannot open"),
    ("mtllib missing.mtl\n", None, "cannot open material library"),
])
def test_errors(tmp_path, obj, mtl, msg):
    with pytest.raises(native.TracerError, match=None) as e:
        _load(tmp_path, obj, mtl)
    assert msg in str(e.value)


def test_missing_file(tmp_path):
    with pytest.raises(native.TracerError) as e:
        native.obj_scene(tmp_path / "nope.obj", CAM)
    assert "cannot open" in str(e.value)


def test_textured_room_loads_and_renders_on_the_oracle(tmp_path):
    """The textured room: both map_Kd materials get their own scaled copy of
    the image, per-corner texture coordinates reach the scene, and the
    oracle's image differs from the untextured room's."""
    from oracle import refcpu
    from tests import objfiles
    from tracer_amd import native
    rng = np.random.default_rng(12)
    checker = np.kron((np.indices((4, 4)).sum(0) % 2)[..., None] * 200 + 30, np.ones((4, 4, 3))).astype(np.uint8)
    checker = (checker * rng.uniform(0.6, 1.0, (16, 16, 3))).astype(np.uint8)
    native.write_png(tmp_path / "checker.png", checker)
    cam = objfiles.camera((0.0, 0.0, 3.6), (0.0, 0.0, 0.0))
    obj, mtl = objfiles.box_room_obj(texture="checker.png")
    s = native.obj_scene(objfiles.write(tmp_path, "room.obj", obj, mtl), cam, outside_fog=False)
    assert len(s.textures) == 2 and s.uvs is not None
    assert sorted(m.base_color_texture for m in s.materials if m.base_color_texture) == [1, 2]
    obj0, mtl0 = objfiles.box_room_obj()
    s0 = native.obj_scene(objfiles.write(tmp_path / "plain", "room.obj", obj0, mtl0), cam, outside_fog=False)
    imgs = []
    for sc in (s, s0):
        sums, _, _ = refcpu.RefScene(sc, 3).render(sc.render_params(width=16, height=16, spp=16, seed=1, max_depth=4),
                                                   quirks=False, nee=0, threads=4)
        imgs.append(sums[..., :3] / sums[..., 3:])
    assert np.isfinite(imgs[0]).all()
    assert np.abs(imgs[0] - imgs[1]).mean() > 0.02 * imgs[1].mean()
