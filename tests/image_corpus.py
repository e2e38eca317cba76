"""Seeded image files for the imread tests (tests/test_image_formats.py)
and their golden fixtures (tests/golden/make_image_fixtures.py): grey JPEG
(baseline, progressive, restart intervals, 16-bit and extreme quantisation
tables, sampling factors, damaged streams), PGM/PPM, TGA (raw, RLE,
colour-mapped, both row orders), PNG variants and the colour-only formats,
written with PIL or byte by byte.  Returns {name: file bytes}."""
from __future__ import annotations

import io
import struct

import numpy as np


def _pil():
    from PIL import Image

    return Image


def _jpeg(a, **kw) -> bytes:
    b = io.BytesIO()
    _pil().fromarray(a).save(b, format="JPEG", **kw)
    return b.getvalue()


def _save(a, fmt, mode=None, **kw) -> bytes:
    b = io.BytesIO()
    _pil().fromarray(a, mode).save(b, format=fmt, **kw) if mode else _pil().fromarray(a).save(b, format=fmt, **kw)
    return b.getvalue()


def tga(type_, w, h, bpp, data, cmap=0, pal=b"", pal_bits=0, pal_start=0, desc=0, ident=b"") -> bytes:
    """A TGA file: the 18-byte header, the image ID, pal_start skipped bytes
    and the palette (colour-mapped images), then the pixel data."""
    n_pal = len(pal) // max(1, pal_bits // 8) if cmap else 0
    hdr = struct.pack("<BBBHHBHHHHBB", len(ident), cmap, type_, pal_start, n_pal, pal_bits, 0, 0, w, h, bpp, desc)
    return hdr + ident + (b"\0" * pal_start if cmap else b"") + pal + data


def tga_rle8(values) -> bytes:
    """TGA run-length packets of 8-bit values: runs of 2..128 equal values
    as repeat packets, the rest as raw packets of up to 128."""
    out, i, n = bytearray(), 0, len(values)
    while i < n:
        j = i
        while j < n and j - i < 128 and values[j] == values[i]:
            j += 1
        if j - i >= 2:
            out += bytes([0x80 | (j - i - 1), values[i]])
            i = j
            continue
        k = i + 1
        while k < n and k - i < 128 and (k + 1 >= n or values[k + 1] != values[k]):
            k += 1
        out += bytes([k - i - 1]) + bytes(values[i:k])
        i = k
    return bytes(out)


def sof_offset(jpeg: bytes) -> int:
    """Offset of the first SOF0/SOF1/SOF2 marker."""
    return min(i for i in (jpeg.find(b"\xff\xc0"), jpeg.find(b"\xff\xc1"), jpeg.find(b"\xff\xc2")) if i >= 0)


def build(seed: int = 5, full: bool = True) -> dict:
    rng = np.random.default_rng(seed)
    h, w = 23, 41
    noise = rng.integers(0, 256, (h, w), dtype=np.uint8)
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = ((np.sin(xx / 4.0) + np.cos(yy / 3.0)) * 60 + 128).astype(np.uint8)
    wide = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    c = {}
    # JPEG
    base = _jpeg(smooth, quality=80)
    c["j_base.jpg"] = base
    c["j_prog_noise.jpg"] = _jpeg(noise, quality=90, progressive=True)
    c["j_rst.jpg"] = _jpeg(noise, quality=85, restart_marker_blocks=3)
    c["j_prog_rst.jpg"] = _jpeg(smooth, quality=85, progressive=True, restart_marker_blocks=2)
    c["j_q16.jpg"] = _jpeg(smooth, qtables=[[(i * 37) % 900 + 1 for i in range(64)]])
    c["j_q1_prog.jpg"] = _jpeg(noise, qtables=[[1] * 64], progressive=True)
    checker = (np.indices((32, 48)).sum(0) % 2 * 255).astype(np.uint8)
    c["j_checker_q1.jpg"] = _jpeg(checker, qtables=[[1] * 64])
    s = sof_offset(base) + 2 + 2 + 1 + 4 + 1 + 1  # the component's sampling-factor byte
    c["j_samp22.jpg"] = base[:s] + b"\x22" + base[s + 1:]
    c["j_trunc.jpg"] = base[: len(base) * 2 // 3]
    c["j_no_eoi.jpg"] = base[:-2]
    c["j_junk_after.jpg"] = base + b"junk\x00\xff\x00trailing"
    c["j_fill_ff.jpg"] = base.replace(b"\xff\xc0", b"\xff\xff\xff\xc0", 1)
    c["j_rgb.jpg"] = _jpeg(np.dstack([noise, smooth, noise]), quality=80)
    # PGM / PPM
    c["p16.pgm"] = b"P5\n41 23\n65535\n" + wide.astype(">u2").tobytes()
    c["p16_1000.pgm"] = b"P5 41 23 1000 " + np.minimum(wide, 1000).astype(">u2").tobytes()
    c["p8_comments.pgm"] = b"P5\r\n# a comment\n41\t# width\n23 #height\n255\r" + noise.tobytes()
    c["p8_trunc.pgm"] = b"P5\n41 23\n255\n" + noise.tobytes()[:-5]
    c["p6.ppm"] = b"P6\n41 23\n255\n" + np.repeat(noise, 3).tobytes()
    c["p2_ascii.pgm"] = b"P2\n2 1\n255\n1 2\n"
    # TGA
    pal = bytes(range(255, -1, -1))
    steps = (smooth // 32 * 32).astype(np.uint8).ravel().tolist()
    c["t3_bottom_up.tga"] = tga(3, w, h, 8, noise.tobytes())
    c["t3_top_down_id.tga"] = tga(3, w, h, 8, noise.tobytes(), desc=0x20, ident=b"an image id")
    c["t2_8bpp.tga"] = tga(2, w, h, 8, noise.tobytes())
    c["t1_pal8.tga"] = tga(1, w, h, 8, noise.tobytes(), cmap=1, pal=pal, pal_bits=8)
    c["t1_pal8_short.tga"] = tga(1, w, h, 8, noise.tobytes(), cmap=1, pal=pal[:100], pal_bits=8)
    c["t1_pal8_start.tga"] = tga(1, w, h, 8, noise.tobytes(), cmap=1, pal=pal, pal_bits=8, pal_start=3)
    c["t1_idx16.tga"] = tga(1, w, h, 16, noise.astype("<u2").tobytes(), cmap=1, pal=pal, pal_bits=8)
    c["t11_rle.tga"] = tga(11, w, h, 8, tga_rle8(steps))
    c["t9_rle_pal.tga"] = tga(9, w, h, 8, tga_rle8(steps), cmap=1, pal=pal, pal_bits=8, desc=0x20)
    c["t3_16bpp.tga"] = tga(3, w, h, 16, noise.astype("<u2").tobytes())
    c["t2_24bpp.tga"] = tga(2, w, h, 24, np.repeat(noise, 3).tobytes())
    if full:
        # PNG and the colour-only formats
        c["g16.png"] = _save(wide, "PNG")
        c["g1.png"] = _save(noise > 128, "PNG")
        c["la.png"] = _save(np.dstack([noise, smooth]), "PNG", "LA")
        c["trns.png"] = _save(noise, "PNG", transparency=7)
        c["g.bmp"] = _save(noise, "BMP")
        c["rgba.bmp"] = _save(np.dstack([noise, smooth, noise, smooth]), "BMP", "RGBA")
        c["g.gif"] = _save(noise, "GIF")
        c["garbage.bin"] = rng.integers(0, 256, 300, dtype=np.uint8).tobytes()
        c["empty.bin"] = b""
    return c


# Where this library deliberately differs from the reference's decoder:
# pixels that decoder leaves uninitialised (a TGA whose pixel data ends early).
UNDEFINED_IN_REFERENCE = {"t3_trunc.tga"}


def truncated_tga():
    rng = np.random.default_rng(9)
    a = rng.integers(0, 256, (23, 41), dtype=np.uint8)
    return a, tga(3, 41, 23, 8, a.tobytes()[:500], desc=0x20)
