"""Image outputs and post processing of the facade (write_image_to_hdr,
write_image_to_png, create_gamma_corrector, create_aces_tone_mapper; the
reference writes with stb_image_write).  Pinned by the formats' definitions:
an RGBE pixel keeps 8 mantissa bits of its largest channel (values such as
1, 0.5 and 3.125 round-trip exactly through this build's reader); the PNG is
parsed here with zlib (chunk CRCs, IHDR fields, filter-0 rows equal to the
input bytes); the ACES curve and gamma have closed forms."""
import struct
import zlib

import numpy as np
import pytest

from tracer_amd import native


@pytest.mark.parametrize("w,h", [(8, 3), (37, 5), (5, 4)])
def test_hdr_round_trip(tmp_path, w, h):
    rng = np.random.default_rng(w)
    img = rng.uniform(0, 4, (h, w, 3)).astype(np.float32)
    img[0] = np.float32([1.0, 0.5, 3.125])      # exactly representable: a run-length row
    img[1, : w // 2] = 0.0                       # black run
    p = tmp_path / "x.hdr"
    native.write_hdr(p, img)
    back = native.read_hdr(p)
    assert back.shape == img.shape
    assert np.array_equal(back[0], img[0]) and np.array_equal(back[1, : w // 2], img[1, : w // 2])
    # every channel within one 8-bit step of the pixel's largest channel
    step = img.max(axis=2, keepdims=True) / 128
    assert np.all(np.abs(back - img) <= step + 1e-12)


def test_hdr_writer_clamps_what_rgbe_cannot_hold(tmp_path):
    """Infinite, huge, NaN and negative channels: the largest RGBE value
    (255/256 * 2^127, exponent byte 255) for the first two, 0 for the others."""
    big = np.float32(np.ldexp(255.0 / 256.0, 127))
    img = np.array([[[np.inf, 1.0, 0.0], [np.finfo(np.float32).max, 0.0, 0.0], [np.nan, 2.0, -1.0],
                     [big, big, big]]], np.float32)
    p = tmp_path / "c.hdr"
    native.write_hdr(p, img)
    back = native.read_hdr(p)
    assert np.isfinite(back).all()
    assert back[0, 0, 0] >= 0.99 * big and back[0, 1, 0] >= 0.99 * big and back[0, 3].min() >= 0.99 * big
    assert back[0, 0, 1] <= 2.0 ** 120 and back[0, 2, 0] == 0.0 and back[0, 2, 2] == 0.0
    assert abs(back[0, 2, 1] - 2.0) <= 2.0 / 128


def test_scene_with_non_finite_vertices_is_rejected():
    import ctypes as C
    from tracer_amd import abi
    s = native.config_scene(1)
    lib = native.hip_lib()
    lib.tr_bvh4_export.argtypes = [C.POINTER(abi.TrSceneDesc), C.POINTER(abi.TrBvhOpts), C.c_void_p, C.c_int64,
                                   C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    n, depth = C.c_int64(), C.c_int32()
    opts = abi.TrBvhOpts(4, 1)
    d = s.desc()
    assert lib.tr_bvh4_export(C.byref(d), C.byref(opts), None, 0, None, C.byref(n), C.byref(depth)) == abi.TR_OK
    for field, value in (("positions", np.nan), ("positions", np.inf), ("normals", np.nan)):
        bad = native.config_scene(1)
        arr = getattr(bad, field)
        if arr is None:
            bad.normals = np.zeros_like(bad.positions) + np.float32([0, 1, 0] * (bad.positions.size // 3))
            arr = bad.normals
        arr.reshape(-1)[7] = value
        d = bad.desc()
        assert lib.tr_bvh4_export(C.byref(d), C.byref(opts), None, 0, None, C.byref(n), C.byref(depth)) == \
            abi.TR_ERR_INVALID, field


def _png_chunks(data: bytes):
    assert data[:8] == bytes([137, 80, 78, 71, 13, 10, 26, 10])
    pos, out = 8, []
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        out.append((kind, body))
        pos += 12 + n
    return out


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (7, 11, 3), dtype=np.uint8)
    p = tmp_path / "x.png"
    native.write_png(p, img)
    chunks = _png_chunks(p.read_bytes())
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (w, h, depth, ctype, comp, filt, inter) == (11, 7, 8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[1][1]), np.uint8).reshape(7, 1 + 11 * 3)
    assert np.all(raw[:, 0] == 0)
    assert np.array_equal(raw[:, 1:].reshape(7, 11, 3), img)


def test_aces_and_gamma_closed_forms():
    x = np.float32([[[0.0, 1.0, 0.25], [100.0, 0.5, 2.0]]])
    a = native.postprocess(x, aces_exposure=1.0)
    curve = lambda v: np.clip(v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14), 0, 1)
    assert a[0, 0, 0] == 0.0 and a[0, 1, 0] == 1.0            # 0 stays 0; large values clamp to 1
    assert abs(a[0, 0, 1] - 2.54 / 3.16) < 1e-6                  # curve(1) = (2.51 + 0.03) / (2.43 + 0.59 + 0.14)
    assert np.allclose(a, curve(x.astype(np.float64)), rtol=1e-6, atol=1e-7)
    e = native.postprocess(x, aces_exposure=2.0)
    assert np.allclose(e, curve(2.0 * x.astype(np.float64)), rtol=1e-6, atol=1e-7)
    g = native.postprocess(x, gamma=1 / 2.2)
    assert np.allclose(g, x.astype(np.float64) ** (1 / 2.2), rtol=1e-6)
    # gamma <= 0 and exposure < 0 leave the image unchanged
    assert np.array_equal(native.postprocess(x), x)


def _png_bytes(img: np.ndarray, filters, ctype=2, palette=None) -> bytes:
    """A PNG built here: rows filtered with the given filter types (0..4)."""
    h, w = img.shape[:2]
    ch = img.shape[2] if img.ndim == 3 else 1
    rows = img.reshape(h, w * ch).astype(np.int64)
    out = bytearray()
    prev = np.zeros(w * ch, np.int64)
    for y in range(h):
        cur = rows[y]
        a = np.concatenate([np.zeros(ch, np.int64), cur[:-ch]])
        c = np.concatenate([np.zeros(ch, np.int64), prev[:-ch]])
        f = filters[y % len(filters)]
        if f == 0: pred = np.zeros_like(cur)
        elif f == 1: pred = a
        elif f == 2: pred = prev
        elif f == 3: pred = (a + prev) // 2
        else:
            p = a + prev - c
            pa, pb, pc = np.abs(p - a), np.abs(p - prev), np.abs(p - c)
            pred = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, prev, c))
        out.append(f)
        out.extend(((cur - pred) & 255).astype(np.uint8).tobytes())
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)
    data = bytes([137, 80, 78, 71, 13, 10, 26, 10]) + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        data += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    return data + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b"")


@pytest.mark.parametrize("filters", [[0], [1], [2], [3], [4], [0, 1, 2, 3, 4]])
def test_png_reader_filters(tmp_path, filters):
    """Each filter type decodes to the original bytes; rows come back bottom
    to top (the reference flips 8-bit images on load)."""
    rng = np.random.default_rng(len(filters) + filters[0])
    img = rng.integers(0, 256, (6, 9, 3), dtype=np.uint8)
    p = tmp_path / "f.png"
    p.write_bytes(_png_bytes(img, filters))
    assert np.array_equal(native.read_png(p), img[::-1])


def test_png_reader_grey_palette_alpha(tmp_path):
    rng = np.random.default_rng(9)
    grey = rng.integers(0, 256, (4, 5), dtype=np.uint8)
    p = tmp_path / "g.png"
    p.write_bytes(_png_bytes(grey, [4], ctype=0))
    assert np.array_equal(native.read_png(p), np.repeat(grey[::-1, :, None], 3, axis=2))
    pal = rng.integers(0, 256, (16, 3))
    idx = rng.integers(0, 16, (4, 5), dtype=np.uint8)
    p.write_bytes(_png_bytes(idx, [0], ctype=3, palette=pal))
    assert np.array_equal(native.read_png(p), pal[idx][::-1].astype(np.uint8))
    rgba = rng.integers(0, 256, (3, 4, 4), dtype=np.uint8)
    p.write_bytes(_png_bytes(rgba, [1], ctype=6))
    assert np.array_equal(native.read_png(p), rgba[::-1, :, :3])
    ga = rng.integers(0, 256, (3, 4, 2), dtype=np.uint8)
    p.write_bytes(_png_bytes(ga, [0], ctype=4))
    with pytest.raises(native.TracerError):
        native.read_png(p)


def test_png_writer_reader_round_trip(tmp_path):
    rng = np.random.default_rng(10)
    img = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    p = tmp_path / "w.png"
    native.write_png(p, img)
    assert np.array_equal(native.read_png(p), img[::-1])


def test_texture_interpolation_q29():
    """Fixed mode interpolates texel / 255 in floating point; the reference
    (Q29) truncates every weighted 8-bit term: between texels 0 and 255 at the
    midpoint it gives 127 / 255 where the exact value is 1/2."""
    from oracle import refcpu
    tex = np.zeros((1, 2, 3), np.uint8)
    tex[0, 1] = 255
    uv = np.float32([[0.5, 0.0], [0.25, 0.0], [1.0 / 3.0, 0.0]])
    fixed = refcpu.tex_eval8(tex, uv, quirks=False)
    ref = refcpu.tex_eval8(tex, uv, quirks=True)
    assert np.allclose(fixed[:, 0], [0.5, 0.25, 1 / 3], atol=1e-6)
    assert np.array_equal(ref[:, 0] * 255, np.float32([127, 63, 85]))


def test_obj_texture_is_decoded_and_scaled(tmp_path):
    """An MTL map_Kd is read next to the MTL file, decoded to linear as
    uint8(255 (c / 255)^2.2) and scaled by Kd as uint8(c * Kd)
    (image_file.cpp:99-130)."""
    from tests import objfiles
    rng = np.random.default_rng(11)
    img = rng.integers(0, 256, (4, 6, 3), dtype=np.uint8)
    native.write_png(tmp_path / "t.png", img)
    obj = "mtllib scene.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nusemtl t\nf 1/1 2/2 3/3\n"
    s = native.obj_scene(objfiles.write(tmp_path, "m.obj", obj, "newmtl t\nKd 0.8 0.5 1.0\nmap_Kd t.png\n"),
                         objfiles.camera((0, 0, 3), (0, 0, 0)), outside_fog=False)
    assert len(s.textures) == 1 and s.materials[0].base_color_texture == 1
    lin = (255 * (img[::-1].astype(np.float32) / 255) ** np.float32(2.2)).astype(np.uint8)
    want = (lin.astype(np.float32) * np.float32([0.8, 0.5, 1.0])).astype(np.uint8)
    got = s.textures[0]
    assert np.max(np.abs(got.astype(int) - want.astype(int))) <= 1   # float pow rounding at integer edges
    assert np.mean(got == want) > 0.99
