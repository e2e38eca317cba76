"""Extracts the reference's inferno table (dt/src/inferno.cpp:4-261, 256
rgb8 triples) into inferno_table.json, as data.  Run in a container where
the reference tree is present:

  python tests/golden/make_inferno_fixture.py /root/reference
"""
import json
import os
import re
import sys


def main(ref_root: str) -> None:
    src = open(os.path.join(ref_root, "dt", "src", "inferno.cpp")).read()
    rows = [[int(a), int(b), int(c)] for a, b, c in re.findall(r"rgb8\{(\d+),\s*(\d+),\s*(\d+)\}", src)]
    assert len(rows) == 256, len(rows)
    out = {"provenance": "dt/src/inferno.cpp:4-261 of baptisteesteban/distance_transform_benchmarks (MPL-2.0)",
           "rgb": rows}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "inferno_table.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
