"""Film output: the reference's 8-bit conversion and the PNG file (host code, no GPU).

The conversion restates Spectrum::gamma_correct + to_rgb (src/spectrum.rs:13-30): sqrt, clamp to
[0, 1] with NaN-ignoring max / min, x255, truncation. The PNG is decoded here with zlib and its
chunk CRCs recomputed independently, so a wrong CRC, length, filter byte or Adler-32 fails.
"""
import struct
import zlib

import numpy as np
import pytest

import pbrt_amd


def reference_rgba8(film, tonemap=False):
    f = np.asarray(film, np.float64)
    with np.errstate(invalid="ignore"):
        if tonemap:
            f = f / (1.0 + f)
        g = np.sqrt(f)
    g = np.where(np.isnan(g), 0.0, g)  # f64::max(0.0, NaN) is 0.0
    rgb = (np.minimum(1.0, np.maximum(0.0, g)) * 255.0).astype(np.uint8)
    return np.concatenate([rgb, np.full(rgb.shape[:2] + (1,), 255, np.uint8)], axis=2)


def decode_png(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        chunks.append((kind, body))
        pos += 12 + n
    assert [k for k, _ in chunks] == [b"IHDR", b"IDAT", b"IEND"]
    w, h, depth, color, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, color, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(chunks[1][1])  # checks the Adler-32 too
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)  # filter type None on every scanline
    return rows[:, 1:].reshape(h, w, 4)


def sample_film(rng, h=37, w=53):
    f = rng.uniform(-0.1, 1.5, (h, w, 3))
    f[0, 0] = [0.0, 1.0, 0.25]             # exact points of the gamma: 0, 255, 127
    f[0, 1] = [np.nan, np.inf, -np.inf]     # NaN and -inf map to 0, +inf to 255
    f[0, 2] = [(k / 255.0) ** 2 for k in (1, 128, 254)]
    return f


def test_rgba8_matches_the_reference_conversion():
    f = sample_film(np.random.default_rng(3))
    got = pbrt_amd.film_to_rgba8(f)
    np.testing.assert_array_equal(got, reference_rgba8(f))
    assert list(got[0, 0]) == [0, 255, 127, 255]
    assert list(got[0, 1, :3]) == [0, 255, 0]


def test_tonemap_option():
    f = np.full((2, 2, 3), 1.0)
    got = pbrt_amd.film_to_rgba8(f, tonemap=True)
    assert np.all(got[..., :3] == int(np.sqrt(0.5) * 255.0))  # 1 / (1 + 1) = 0.5
    f = sample_film(np.random.default_rng(4))
    np.testing.assert_array_equal(pbrt_amd.film_to_rgba8(f, tonemap=True), reference_rgba8(f, tonemap=True))


@pytest.mark.parametrize("shape", [(1, 1), (3, 200), (300, 257)])  # 300 x 257 spans several stored blocks
def test_png_decodes_to_the_pixels_and_is_deterministic(tmp_path, shape):
    f = np.random.default_rng(5).uniform(0.0, 1.2, shape + (3,))
    a, b = tmp_path / "a.png", tmp_path / "b.png"
    pbrt_amd.write_png(f, a)
    pbrt_amd.write_png(f, b)
    data = a.read_bytes()
    assert data == b.read_bytes()
    np.testing.assert_array_equal(decode_png(data), pbrt_amd.film_to_rgba8(f))


def test_film_errors(tmp_path):
    with pytest.raises(pbrt_amd.PbrtError):
        pbrt_amd.film_to_rgba8(np.zeros((0, 4, 3)))
    with pytest.raises(pbrt_amd.PbrtError):
        pbrt_amd.write_png(np.zeros((2, 2, 3)), tmp_path / "no_such_dir" / "x.png")
