"""Writes tests/golden/images/: a few small image files from
tests/image_corpus.py and, in expected.npz, what the reference's own
decoder returns for each (stb_image 2.30 as vendored in the reference,
built into oracle/_ref by oracle/Makefile; dt/src/imread.cpp:12 calls
stbi_load(filename, &w, &h, &n, 0)).  Per file the archive holds the (H, W)
uint8 pixels of a one-component image, or a 0-d array with the component
count stb reports otherwise.

  python tests/golden/make_image_fixtures.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

GOLDEN = ["j_base.jpg", "j_prog_noise.jpg", "j_rst.jpg", "j_prog_rst.jpg", "j_q16.jpg", "j_samp22.jpg",
          "j_trunc.jpg", "j_rgb.jpg", "p16.pgm", "p8_comments.pgm", "t3_bottom_up.tga", "t1_idx16.tga",
          "t9_rle_pal.tga", "t3_16bpp.tga"]


def main() -> None:
    import image_corpus
    from oracle import oracle

    oracle.build()
    if not oracle.stb_available():
        raise SystemExit("the reference's stb_image is not built (oracle/_ref): the reference tree is needed")
    files = image_corpus.build(full=False)
    out = os.path.join(HERE, "images")
    os.makedirs(out, exist_ok=True)
    expected = {}
    for name in GOLDEN:
        path = os.path.join(out, name)
        with open(path, "wb") as f:
            f.write(files[name])
        px, n = oracle.stb_load(path)
        expected[name] = px[:, :, 0] if n == 1 else np.array(n)
    np.savez_compressed(os.path.join(out, "expected.npz"), **expected)
    print(f"wrote {len(GOLDEN)} files and expected.npz to {out}")


if __name__ == "__main__":
    main()
