"""Write tests/golden/images/: small TGA and PNG files in the variants the
texture loader handles, plus images.npz with the RGBA8 pixels Pillow
decodes from each (the expected values).

TGA files are written by Pillow (raw / RLE, bottom-left / top-left origin,
32 / 24 / 8 bits).  Pillow cannot write 16-bit RGBA or filtered PNGs on
demand, so the PNGs are assembled here from zlib and struct (every scanline
filter type in turn, 8 and 16 bits, grey / grey+alpha / RGB / RGBA /
palette) and then decoded by Pillow: the expected values come from
Pillow's decoder, not from this writer.
"""
import os
import struct
import zlib

import numpy as np
from PIL import Image

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "images")


def chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def filter_rows(raw: np.ndarray, bpp: int) -> bytes:
    """raw: [h, stride] uint8; row y uses filter y % 5."""
    h, stride = raw.shape
    out = bytearray()
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        f = y % 5
        cur = raw[y].astype(np.int64)
        row = np.zeros(stride, np.int64)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) >> 1, paeth(a, b, c)][f]
            row[i] = (cur[i] - pred) & 0xFF
        out.append(f)
        out += bytes(row.astype(np.uint8))
        prev = cur
    return bytes(out)


def write_png(path, samples: np.ndarray, ctype: int, depth: int, palette=None, trns=None):
    """samples: [h, w, channels] (uint8 or uint16)."""
    h, w = samples.shape[:2]
    if depth == 16:
        raw = samples.astype(">u2").view(np.uint8).reshape(h, -1)
    else:
        raw = samples.astype(np.uint8).reshape(h, -1)
    bpp = samples.shape[2] * depth // 8
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if palette is not None:
        data += chunk(b"PLTE", bytes(palette.astype(np.uint8).reshape(-1)))
    if trns is not None:
        data += chunk(b"tRNS", bytes(trns.astype(np.uint8)))
    comp = zlib.compress(filter_rows(raw, bpp), 9)
    data += chunk(b"IDAT", comp[: len(comp) // 2]) + chunk(b"IDAT", comp[len(comp) // 2:]) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def main():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(7)
    expected = {}
    # TGA: runs of equal pixels so RLE packets of both kinds occur
    base = rng.integers(0, 256, (13, 11, 4), dtype=np.uint8)
    base[2:5, :, :] = base[2, 0]
    for name, mode, kw in [("tga_rle32_bl", "RGBA", {"compression": "tga_rle"}),
                           ("tga_raw32_tl", "RGBA", {"orientation": 1}),
                           ("tga_rle24_bl", "RGB", {"compression": "tga_rle"}),
                           ("tga_raw8_bl", "L", {})]:
        im = Image.fromarray(base).convert(mode)
        path = os.path.join(OUT, name + ".tga")
        im.save(path, **kw)
        expected[name] = np.asarray(Image.open(path).convert("RGBA"))
    # PNG
    h, w = 9, 7
    rgba16 = rng.integers(0, 65536, (h, w, 4), dtype=np.uint16)
    cases = {
        "png_rgba8": (rng.integers(0, 256, (h, w, 4), dtype=np.uint8), 6, 8, None, None),
        "png_rgba16": (rgba16, 6, 16, None, None),
        "png_rgb8": (rng.integers(0, 256, (h, w, 3), dtype=np.uint8), 2, 8, None, None),
        "png_grey8": (rng.integers(0, 256, (h, w, 1), dtype=np.uint8), 0, 8, None, None),
        "png_greyalpha8": (rng.integers(0, 256, (h, w, 2), dtype=np.uint8), 4, 8, None, None),
        "png_palette8": (rng.integers(0, 6, (h, w, 1), dtype=np.uint8), 3, 8,
                         rng.integers(0, 256, (6, 3), dtype=np.uint8), np.array([255, 0, 128])),
    }
    for name, (samples, ctype, depth, pal, trns) in cases.items():
        path = os.path.join(OUT, name + ".png")
        write_png(path, samples, ctype, depth, pal, trns)
        expected[name] = np.asarray(Image.open(path).convert("RGBA"))
    np.savez(os.path.join(OUT, "images.npz"), **expected)
    print("wrote", sorted(expected))


if __name__ == "__main__":
    main()
