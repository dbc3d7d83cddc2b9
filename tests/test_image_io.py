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
