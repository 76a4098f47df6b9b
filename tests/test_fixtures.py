"""The texture loaders (meteoros_amd/csrc/fixtures.cpp) against Pillow.

Committed fixtures (tests/golden/images/, made by make_image_fixtures.py):
every TGA variant the loader takes (raw / RLE, bottom-left / top-left,
32 / 24 / 8 bits) and PNGs with every scanline filter at 8 and 16 bits in
all five colour types, each compared byte for byte with Pillow's RGBA
decode.  When the reference's texture folder is present (this container,
not the GPU box) its shipped volumes and maps are loaded too and must equal
Pillow's decode and the statistics in tests/golden/texture_stats.json.
"""
import json
import os

import numpy as np
import pytest

import meteoros_amd as M

HERE = os.path.dirname(os.path.abspath(__file__))
IMG = os.path.join(HERE, "golden", "images")
SHIPPED = "/root/reference/src/CloudScapes/textures/CloudTextures"


def _expected():
    with np.load(os.path.join(IMG, "images.npz")) as d:
        return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name", sorted(_expected()))
def test_loader_matches_pillow_on_fixtures(name):
    exp = _expected()[name]
    ext = ".tga" if name.startswith("tga") else ".png"
    got = M.load_image_rgba8(os.path.join(IMG, name + ext))
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


def test_sixteen_bit_samples_keep_their_high_byte():
    # the 16-bit fixture's expected bytes are Pillow's; stbi_load(..., 4)
    # also keeps the high byte, so both agree with a direct read of the file
    import zlib

    raw = open(os.path.join(IMG, "png_rgba16.png"), "rb").read()
    assert raw[24] == 16 and raw[25] == 6  # bit depth, colour type
    got = M.load_image_rgba8(os.path.join(IMG, "png_rgba16.png"))
    assert got.dtype == np.uint8 and got.shape == (9, 7, 4)
    assert zlib.crc32(got.tobytes()) == zlib.crc32(_expected()["png_rgba16"].tobytes())


def test_loader_errors(tmp_path):
    with pytest.raises(M.MtcError):
        M.load_image_rgba8(str(tmp_path / "missing.png"))
    bad = tmp_path / "bad.png"
    bad.write_bytes(b"\x89PNG\r\n\x1a\n" + b"\x00" * 20)
    with pytest.raises(M.MtcError):
        M.load_image_rgba8(str(bad))
    trunc = tmp_path / "t.tga"
    trunc.write_bytes(open(os.path.join(IMG, "tga_rle32_bl.tga"), "rb").read()[:60])
    with pytest.raises(M.MtcError):
        M.load_image_rgba8(str(trunc))
    with pytest.raises(M.MtcError):
        M.load_tga_slices(str(tmp_path), "none", 4)


def test_volume_from_slices_layout(tmp_path):
    # slice z <- "<base>(z+1).tga", x fastest, then y, then z
    from PIL import Image

    n = 4
    rng = np.random.default_rng(2)
    vol = rng.integers(0, 256, (n, n, n, 4), dtype=np.uint8)
    for z in range(n):
        Image.fromarray(vol[z], "RGBA").save(str(tmp_path / f"Vol({z + 1}).tga"), compression="tga_rle")
    got = M.load_tga_slices(str(tmp_path), "Vol", n)
    assert np.array_equal(got, vol.reshape(-1))


@pytest.mark.skipif(not os.path.isdir(SHIPPED), reason="the reference's texture folder is not on this machine")
def test_shipped_textures_load_as_pillow_reads_them():
    from PIL import Image

    lf, hf, curl, weather = M.load_shipped_volumes(SHIPPED)
    assert lf.size == 128 ** 3 * 4 and hf.size == 32 ** 3 * 4
    assert curl.size == 128 * 128 * 4 and weather.size == 512 * 512 * 4
    for z in (0, 63, 127):
        ref = np.asarray(Image.open(f"{SHIPPED}/LowFrequency/LowFrequency({z + 1}).tga").convert("RGBA"))
        assert np.array_equal(lf.reshape(128, 128, 128, 4)[z], ref)
    ref = np.asarray(Image.open(f"{SHIPPED}/weatherMap.png").convert("RGBA"))
    assert np.array_equal(weather.reshape(512, 512, 4), ref)
    stats = json.load(open(os.path.join(HERE, "golden", "texture_stats.json")))
    for name, arr in (("lf", lf), ("hf", hf), ("curl", curl), ("weather", weather)):
        a = arr.reshape(-1, 4).astype(np.float64) / 255.0
        assert np.allclose(a.mean(0), stats[name]["mean"], atol=1e-12), name
        assert np.allclose(a.std(0), stats[name]["std"], atol=1e-12), name


def test_headers_claiming_huge_images_are_refused_before_allocating(tmp_path):
    """Sizes come from the file header: a 65535^2 TGA of 18 bytes, an RLE TGA
    whose few packets cannot cover its claimed size, and a 65536^2 PNG with a
    tiny IDAT are refused with an error status (no allocation of the claimed
    size, no exception across the C ABI)."""
    import struct
    import zlib

    from meteoros_amd import _abi as A

    def tga(w, h, type_, body=b""):
        return bytes([0, 0, type_]) + b"\x00" * 9 + struct.pack("<HHBB", w, h, 32, 0x28) + body

    cases = {
        "raw_huge.tga": tga(65535, 65535, 2),
        "raw_short.tga": tga(64, 64, 2, b"\x01" * 100),
        "rle_huge.tga": tga(65535, 65535, 10, b"\xff" + b"\x00" * 4),
    }

    def chunk(kind, data):
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data))

    ihdr = struct.pack(">IIBBBBB", 65536, 65536, 8, 6, 0, 0, 0)
    cases["huge.png"] = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(b"\x00" * 64))
                         + chunk(b"IEND", b""))
    ihdr = struct.pack(">IIBBBBB", 16384, 16384, 8, 6, 0, 0, 0)
    cases["too_many_pixels.png"] = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr)
                                    + chunk(b"IDAT", zlib.compress(b"\x00" * 64)) + chunk(b"IEND", b""))
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        with pytest.raises(M.MtcError) as e:
            M.load_image_rgba8(str(p))
        assert e.value.code == A.MTC_ERR_INVALID_ARG, name
