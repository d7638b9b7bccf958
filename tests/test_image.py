"""Post-process and image files on the host (no GPU): the SDR conversion against the oracle's
restatement (bit-exact) and against double-precision mathematics, the deterministic exp / pow against
libm, BMP bytes against stb_image_write v1.14 (tests/golden/stb_bmp.npz, tests/golden/make_image_fixture.py),
and PNG / EXR files decoded by independent readers written here from the PNG and OpenEXR specifications.

The reference's PNG and EXR writers (fpng, tinyexr) are submodules absent from the reference tree, so
those files are pinned by what they decode to, not by their bytes.
"""
import ctypes as C
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden" / "stb_bmp.npz"


def test_exp_and_pow_against_libm():
    from oracle_lib import oracle

    o = oracle()
    xs = np.concatenate([np.linspace(-103.0, 88.7, 20001, dtype=np.float32),
                         np.array([0.0, -0.0, 1e-30, -1e-30, 0.3466, -0.3466], np.float32)])
    got = np.array([o.oracle_expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    normal = want > 1.2e-38
    assert np.max(np.abs(got[normal] - want[normal]) / ulp[normal]) <= 1.5
    assert np.all(np.abs(got[~normal] - want[~normal]) <= 1.5e-45)
    assert o.oracle_expf(89.0) == np.inf and o.oracle_expf(-104.0) == 0.0 and np.isnan(o.oracle_expf(np.nan))
    # pow on the tone curve's and the sRGB curve's exponents: relative error <= (1 + |y ln x|) 2^-22
    for y in (np.float32(1.0) / np.float32(0.6), np.float32(1.0 / 2.4)):
        xs = np.concatenate([np.geomspace(1e-6, 1e3, 20000).astype(np.float32), np.float32([1.0, 2.0])])
        got = np.array([o.oracle_powf(float(x), float(y)) for x in xs], np.float64)
        want = np.power(xs.astype(np.float64), float(y))
        bound = (1.0 + np.abs(float(y) * np.log(xs.astype(np.float64)))) * 2.0 ** -22
        assert np.all(np.abs(got - want) / want <= bound)
    assert o.oracle_powf(0.0, 1.5) == 0.0 and o.oracle_powf(1.0, 7.0) == 1.0 and o.oracle_powf(5.0, 0.0) == 1.0
    assert np.isnan(o.oracle_powf(-1.0, 0.5)) and o.oracle_powf(np.inf, 0.5) == np.inf


def random_image(rng, h, w, c, specials=True):
    img = rng.gamma(0.7, 0.6, (h, w, c)).astype(np.float32)
    if specials:
        flat = img.reshape(-1)
        idx = rng.choice(flat.size, size=min(flat.size, 40), replace=False)
        flat[idx] = np.float32([np.nan, np.inf, -np.inf, -0.5, 0.0, -0.0, 1e-8, 1e30] * 5)[: idx.size]
    return img


CONFIGS = [
    dict(alpha_override=1.0, brightness_scale=1.0, apply_tone_map=False, apply_srgb_gamma=True, flip_y=False),
    dict(alpha_override=-1.0, brightness_scale=2.5, apply_tone_map=True, apply_srgb_gamma=True, flip_y=True),
    dict(alpha_override=0.5, brightness_scale=0.3, apply_tone_map=True, apply_srgb_gamma=False, flip_y=False),
    dict(alpha_override=-1.0, brightness_scale=1.0, apply_tone_map=False, apply_srgb_gamma=False, flip_y=True),
]


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("channels", [1, 3, 4])
def test_sdr_convert_matches_oracle_bit_exact_and_the_mathematics(cfg, channels):
    from rtcamp8_amd import image
    from oracle_lib import oracle

    rng = np.random.default_rng(channels * 10 + CONFIGS.index(cfg))
    img = random_image(rng, 37, 53, channels)
    sc = image.sdr_config(**cfg)
    got = image.sdr_convert(img, sc)
    for f64 in (0, 1):
        want = np.empty_like(got)
        oracle().oracle_sdr_convert(img.ctypes.data, 53, 37, channels, C.byref(sc), want.ctypes.data, f64)
        if f64 == 0:
            assert np.array_equal(got, want)
        else:  # double-precision evaluation: the same codes except at rounding boundaries
            a = got.view(np.uint8).astype(np.int32)
            b = want.view(np.uint8).astype(np.int32)
            assert np.max(np.abs(a - b)) <= 1 and np.mean(a != b) < 2e-3


def test_sdr_convert_defined_cases():
    """Quantisation is min(trunc(255 v), 255); NaN and negatives give 0; the batch settings write alpha 255;
    flip_y reverses rows; the tone curve maps luminance L to 1 - exp(-L)."""
    from rtcamp8_amd import image

    v = np.float32([[0.0, 1.0 / 255.0, 0.5, 1.0, 2.0, np.nan, -1.0, np.inf]]).T[:, :, None].reshape(1, 8, 1)
    raw = image.sdr_convert(np.repeat(v, 3, axis=2), image.sdr_config(apply_srgb_gamma=False, alpha_override=-1))
    assert [int(p & 255) for p in raw[0]] == [0, 1, 127, 255, 255, 0, 0, 255]
    assert all(int(p >> 24) == 255 for p in raw[0])
    img = np.arange(12, dtype=np.float32).reshape(4, 1, 3) / 12
    a = image.sdr_convert(img, image.sdr_config(flip_y=True))
    b = image.sdr_convert(img[::-1], image.sdr_config(flip_y=False))
    assert np.array_equal(a, b)
    grey = np.full((1, 1, 3), 0.8, np.float32)
    t = image.sdr_convert(grey, image.sdr_config(apply_tone_map=True, apply_srgb_gamma=False))
    lum = 0.8 * (0.2126729 + 0.7151522 + 0.0721750)
    assert int(t[0, 0] & 255) == int(255 * 0.8 * (1 - np.exp(-lum)) / lum)


@pytest.mark.parametrize("name", ["1x1", "2x3", "13x7", "16x16", "33x9"])
def test_bmp_is_stb_image_write_bytes(name):
    from rtcamp8_amd import image

    fx = np.load(GOLDEN)
    assert image.encode_bmp(fx[f"rgba_{name}"]) == fx[f"bmp_{name}"].tobytes()


def decode_png(data: bytes) -> np.ndarray:
    """A PNG reader for 8-bit RGBA, non-interlaced files (PNG specification, sections 5, 9 and 11)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    off, idat, w = 8, b"", None
    while off < len(data):
        (n,) = struct.unpack(">I", data[off:off + 4])
        typ, body = data[off + 4:off + 8], data[off + 8:off + 8 + n]
        (crc,) = struct.unpack(">I", data[off + 8 + n:off + 12 + n])
        assert crc == zlib.crc32(typ + body)
        if typ == b"IHDR":
            w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
        off += 12 + n
    raw = zlib.decompress(idat)
    stride = 4 * w
    out = np.zeros((h, stride), np.uint8)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        up = out[y - 1].astype(np.int32) if y else np.zeros(stride, np.int32)
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - 4] if i >= 4 else 0
            b, c = up[i], (up[i - 4] if i >= 4 else 0)
            if f == 0:
                pred = 0
            elif f == 1:
                pred = a
            elif f == 2:
                pred = b
            elif f == 3:
                pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + pred) & 255
        out[y] = cur
    return out.view("<u4").reshape(h, w)


def test_png_decodes_to_its_pixels():
    from rtcamp8_amd import image

    rng = np.random.default_rng(3)
    smooth = (np.add.outer(np.arange(20), np.arange(31)) * 4).astype(np.uint32) & 255
    for px in (rng.integers(0, 2 ** 32, (17, 23), dtype=np.uint64).astype(np.uint32),
               smooth | (smooth << 8) | (np.uint32(255) << 24),
               np.zeros((1, 1), np.uint32)):
        data = image.encode_png(px)
        assert np.array_equal(decode_png(data), px)


def decode_exr(data: bytes):
    """A reader for single-part scanline OpenEXR files without compression (OpenEXR file layout)."""
    magic, version = struct.unpack("<II", data[:8])
    assert magic == 20000630 and version == 2
    off, attrs = 8, {}
    while data[off] != 0:
        end = data.index(b"\0", off)
        name = data[off:end].decode()
        tend = data.index(b"\0", end + 1)
        typ = data[end + 1:tend].decode()
        (size,) = struct.unpack("<i", data[tend + 1:tend + 5])
        attrs[name] = (typ, data[tend + 5:tend + 5 + size])
        off = tend + 5 + size
    off += 1
    chans, body = [], attrs["channels"][1]
    p = 0
    while body[p] != 0:
        end = body.index(b"\0", p)
        ptype, _, xs, ys = struct.unpack("<iIii", body[end + 1:end + 17])
        chans.append((body[p:end].decode(), ptype, xs, ys))
        p = end + 17
    assert attrs["compression"][1] == b"\0" and attrs["lineOrder"][1] == b"\0"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack(f"<{h}Q", data[off:off + 8 * h])
    planes = {c[0]: np.zeros((h, w), np.float16) for c in chans}
    for y, o in enumerate(offsets):
        yy, size = struct.unpack("<ii", data[o:o + 8])
        assert yy == y and size == 2 * w * len(chans)
        for k, (name, ptype, _, _) in enumerate(chans):
            assert ptype == 1
            planes[name][y] = np.frombuffer(data, "<f2", w, o + 8 + 2 * w * k)
    return [c[0] for c in chans], planes


@pytest.mark.parametrize("channels", [1, 3, 4])
def test_exr_decodes_to_half_rounded_values(channels):
    from rtcamp8_amd import image

    rng = np.random.default_rng(channels)
    img = (rng.standard_normal((9, 14, channels)) * 40).astype(np.float32)
    flat = img.reshape(-1)
    flat[:12] = [65519.0, 65520.0, 1e6, -1e6, 6e-8, 2.98e-8, 2.9802322e-8, 1e-9, -0.0, np.inf, 6.1e-5, 3.0e-5]
    for flip in (False, True):
        names, planes = decode_exr(image.encode_exr(img, brightness_scale=1.5, flip_y=flip))
        assert names == ["A", "B", "G", "R"]
        src = img[::-1] if flip else img
        for c, name in enumerate(["R", "G", "B", "A"]):
            want = (np.float32(1.5) * src[:, :, c]).astype(np.float16) if c < channels else np.zeros(src.shape[:2],
                                                                                                    np.float16)
            assert np.array_equal(planes[name].view(np.uint16), want.view(np.uint16)), name
    _, planes = decode_exr(image.encode_exr(np.float32([[[np.nan]]])))
    assert np.isnan(planes["R"][0, 0])
