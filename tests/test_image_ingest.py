"""The Radiance .hdr reader behind create_texture2d_from_file (facade
image_load.cpp; the reference decodes with stb_image, which it vendors as
dep/stb_image.h).  Pinned by hand-made files whose texels follow from the
format's definition: a pixel (r, g, b, e) is (r, g, b) * 2^(e - 136), 0 when
e = 0; run-length scanlines and flat scanlines decode alike."""
import numpy as np
import pytest

from tracer_amd import native

HEADER = b"#?RADIANCE\n# made by hand\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 8\n"


def _rle_row():
    # R: a run of 8 x 128; G: 8 literals 0, 16, ..., 112; B: a run of 8 x 64; E: a run of 8 x 129
    return bytes([2, 2, 0, 8, 128 + 8, 128, 8, *[16 * k for k in range(8)], 128 + 8, 64, 128 + 8, 129])


def _flat_row():
    return bytes([1, 2, 3, 0]) + bytes([200, 100, 50, 130]) * 7


def test_known_texels(tmp_path):
    p = tmp_path / "t.hdr"
    p.write_bytes(HEADER + _rle_row() + _flat_row())
    img = native.read_hdr(p)
    assert img.shape == (2, 8, 3)
    # row 0: e = 129 -> scale 2^-7
    assert np.array_equal(img[0, :, 0], np.full(8, 1.0, np.float32))
    assert np.array_equal(img[0, :, 1], np.arange(8, dtype=np.float32) / 8)
    assert np.array_equal(img[0, :, 2], np.full(8, 0.5, np.float32))
    # row 1: e = 0 is black, e = 130 -> scale 2^-6
    assert np.array_equal(img[1, 0], np.zeros(3, np.float32))
    assert np.array_equal(img[1, 1:], np.tile(np.float32([3.125, 1.5625, 0.78125]), (7, 1)))


@pytest.mark.parametrize("data,msg", [
    (b"P6\n8 2\n255\n", "not a Radiance picture"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 2 +X 8\n", "unsupported FORMAT"),
    (b"#?RADIANCE\n\n+Y 2 +X 8\n", "orientation"),
    (HEADER + bytes([2, 2, 0, 8, 128 + 9, 1]), "run past the end"),
    (HEADER + bytes(10), "truncated"),
])
def test_errors(tmp_path, data, msg):
    p = tmp_path / "bad.hdr"
    p.write_bytes(data)
    with pytest.raises(native.TracerError) as e:
        native.read_hdr(p)
    assert msg in str(e.value)
