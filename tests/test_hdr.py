"""The RGBE decoder (lucent_decode_hdr) against parseHDR's known answers
(reference test/environment.test.ts:58-98), plus the axis orders and the
malformed inputs parseHDR rejects (src/loaders/hdr.ts:7-93)."""
import numpy as np
import pytest

from lucent_amd import _native
from lucent_amd.assets import decode_hdr


def header(res: str) -> bytes:
    return f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1\n\n{res}\n".encode()


def test_flat_rgbe_pixels():
    # e = 129: scale 2^(129-136) = 1/128; a zero exponent decodes to black.
    w, h, img = decode_hdr(header("-Y 1 +X 2") + bytes([127, 63, 0, 129, 0, 0, 0, 0]))
    assert (w, h) == (2, 1)
    assert img[0, 0].tolist() == [127.5 / 128, 63.5 / 128, 0.5 / 128]
    assert img[0, 1].tolist() == [0.0, 0.0, 0.0]


def test_run_length_encoded_scanline():
    width = 8
    rle = bytes([2, 2, 0, width,
                 128 + 8, 200,
                 4, 1, 2, 3, 4, 128 + 4, 9,
                 128 + 8, 0,
                 128 + 8, 128])
    _, _, img = decode_hdr(header(f"-Y 1 +X {width}") + rle)
    f = 2.0 ** -8
    assert np.all(img[0, :, 0] == 200.5 * f)
    assert img[0, :, 1].tolist() == [(g + 0.5) * f for g in [1, 2, 3, 4, 9, 9, 9, 9]]
    assert np.all(img[0, :, 2] == 0.5 * f)


def test_bottom_up_files_are_flipped():
    top, bottom = [255, 0, 0, 128], [0, 0, 255, 128]
    _, _, img = decode_hdr(header("+Y 2 +X 1") + bytes(bottom + top))
    assert img[0, 0, 0] > 0.5 and img[1, 0, 2] > 0.5


@pytest.mark.parametrize("res", ["-Y 2 +X 3", "+Y 2 +X 3", "-Y 2 -X 3", "+X 3 -Y 2", "-X 3 +Y 2"])
def test_axis_orders(res):
    """Pixel (x, y) of the top-down image holds mantissa 10*y + x + 1 in red."""
    s1, a1, n1, s2, a2, n2 = res[0], res[1], int(res.split()[1]), res.split()[2][0], res.split()[2][1], int(res.split()[3])
    row_y = a1 == "Y"
    width, height = (n2, n1) if row_y else (n1, n2)
    flip_y = (s1 if row_y else s2) == "+"
    flip_x = (s2 if row_y else s1) == "-"
    data = []
    for s in range(n1):
        for i in range(n2):
            x, y = (i, s) if row_y else (s, i)
            if flip_x:
                x = width - 1 - x
            if flip_y:
                y = height - 1 - y
            data += [10 * y + x + 1, 0, 0, 136]
    w, h, img = decode_hdr(header(res) + bytes(data))
    assert (w, h) == (width, height)
    want = np.array([[10 * y + x + 1.5 for x in range(width)] for y in range(height)], np.float32)
    assert np.array_equal(img[..., 0], want)


def test_exponent_range_is_exact():
    _, _, img = decode_hdr(header("-Y 1 +X 2") + bytes([255, 255, 255, 255, 1, 0, 0, 1]))
    assert img[0, 0, 0] == np.float32(255.5 * 2.0 ** 119)
    assert img[0, 1, 0] == np.float32(1.5 * 2.0 ** -135)


@pytest.mark.parametrize("data, msg", [
    (b"P6\n1 1\n255\n", b"Radiance"),
    (header("-Y 2 +X 2") + bytes(4), b"truncated"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n" + bytes(4), b"unsupported format"),
    (header("-Y 1 -Y 1") + bytes(4), b"malformed resolution"),
    (header("-Y 0 +X 4"), b"empty image"),
    (b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe", b"end of header"),
    (header("-Y 1 +X 8") + bytes([2, 2, 0, 9]), b"width mismatch"),
    (header("-Y 1 +X 8") + bytes([2, 2, 0, 8, 128 + 9, 1]), b"run overflows"),
    (header("-Y 1 +X 8") + bytes([2, 2, 0, 8, 0]), b"bad literal run"),
    (header("-Y 1 +X 8") + bytes([2, 2, 0, 8, 128 + 8]), b"truncated run-length"),
])
def test_rejects_malformed_files(data, msg):
    with pytest.raises(_native.LucentError) as e:
        decode_hdr(data)
    assert msg.decode() in str(e.value)
