"""Writes dt_inferno_table.inc: the 256-entry inferno colour map as uint8
RGB triples, floor(255 * c) of matplotlib's published inferno colours (the
quantisation that gives the reference's table, dt/src/inferno.cpp:4-261;
tests/golden/inferno_table.json holds that table and the CPU tests compare).

  python distance_transform_benchmarks_amd/csrc/gen_inferno.py
"""
import os

import numpy as np
from matplotlib import colormaps

HERE = os.path.dirname(os.path.abspath(__file__))


def table() -> np.ndarray:
    rgb = np.asarray(colormaps["inferno"].colors, dtype=np.float64)[:, :3]
    assert rgb.shape == (256, 3)
    return np.floor(rgb * 255.0).astype(np.int64)


def main() -> None:
    t = table()
    lines = ["// Generated by gen_inferno.py: floor(255 * c) of matplotlib's inferno colours, r, g, b per entry."]
    for i in range(0, 256, 4):
        lines.append(" ".join(f"{r:3d}, {g:3d}, {b:3d}," for r, g, b in t[i:i + 4]))
    with open(os.path.join(HERE, "dt_inferno_table.inc"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
