"""Extract the two known-answer vectors of the reference test suite as data.

Source (MPL-2.0, baptisteesteban/distance_transform_benchmarks):
  tests/distance_transform.cpp:14-281   EuclidianFastGeodis  (50x40, lambda=0)
  tests/distance_transform.cpp:283-476  GeodesicFastGeodis   (36x34, lambda=1)

The reference asserts bit-exact equality of its raster GGDT with these
float literals (tests/helpers.hpp:29).  This script parses the literals and
writes them, with the test set-up, to JSON fixtures next to itself.  It reads
the reference tree only when run by hand in the build container; the
fixtures it writes are what the tests load.

Float literals are converted like the C++ initialiser does: decimal ->
double (nearest) -> float (nearest); image literals such as ``122.0`` are
converted to uint8.
"""
from __future__ import annotations

import json
import os
import re
import sys

import numpy as np

REF = os.environ.get("DT_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "tests", "distance_transform.cpp")
HERE = os.path.dirname(os.path.abspath(__file__))


def _array_after(text: str, start: int, name: str):
    i = text.index(name, start)
    j = text.index("{", i)
    k = text.index("};", j)
    body = text[j + 1:k]
    body = re.sub(r"//[^\n]*", "", body)
    vals = [t.strip() for t in body.replace("\n", " ").split(",") if t.strip()]
    return vals, k


def main() -> int:
    text = open(SRC, encoding="utf-8").read()

    # EuclidianFastGeodis: WIDTH 50, HEIGHT 40, zero image, seed rectangle
    # x in [20, 43), y in [20, 30), lambda 0, v default 1e10.
    t0 = text.index("TEST(DistanceTransform, EuclidianFastGeodis)")
    ref0, _ = _array_after(text, t0, "float ref_data[]")
    D0 = np.array([np.float32(float(v)) for v in ref0], np.float32)
    assert D0.size == 50 * 40, D0.size
    fx0 = {
        "provenance": "tests/distance_transform.cpp:14-281 of baptisteesteban/distance_transform_benchmarks (MPL-2.0)",
        "width": 50, "height": 40, "lambda": 0.0, "v": 1e10,
        "image": "zeros", "seed_rect": [20, 20, 43, 30],
        "expected_f32_bits": [int(b) for b in D0.view(np.uint32)],
    }

    t1 = text.index("TEST(DistanceTransform, GeodesicFastGeodis)")
    img1, k = _array_after(text, t1, "std::uint8_t img_data[]")
    ref1, _ = _array_after(text, k, "float ref_data[]")
    I1 = np.array([int(float(v)) for v in img1], np.int64)
    assert I1.size == 36 * 34 and I1.min() >= 0 and I1.max() <= 255
    D1 = np.array([np.float32(float(v)) for v in ref1], np.float32)
    assert D1.size == 36 * 34, D1.size
    fx1 = {
        "provenance": "tests/distance_transform.cpp:283-476 of baptisteesteban/distance_transform_benchmarks (MPL-2.0)",
        "width": 36, "height": 34, "lambda": 1.0, "v": 1e10,
        "image_u8": [int(v) for v in I1], "seed": [18, 17],
        "expected_f32_bits": [int(b) for b in D1.view(np.uint32)],
    }
    for name, fx in (("fastgeodis_euclidean.json", fx0), ("fastgeodis_geodesic.json", fx1)):
        with open(os.path.join(HERE, name), "w", encoding="utf-8") as f:
            json.dump(fx, f, separators=(",", ":"))
        print("wrote", name)
    return 0


if __name__ == "__main__":
    sys.exit(main())
