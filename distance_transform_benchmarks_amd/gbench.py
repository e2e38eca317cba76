"""Benchmark harness with google-benchmark compatible JSON output.

Mirrors benchmarks/distance_transform.cpp of the reference (lines 76-167):
three families DistanceTransformGeos / DistanceTransformBlocks /
DistanceTransformTask (raster, blocks and task schedules), arguments
(image index, lambda index) with lambda in {0, 0.5, 1}, each iteration timed
with a device event pair around the whole allocating call, ``real_time`` in
ms, ``bytes_per_second`` = pixels per second (the reference sets
``SetBytesProcessed`` to the pixel count) and the ``lambda`` / ``size``
counters.  The result file has the reference's ``context`` / ``benchmarks``
layout (results/benchmark_results.json), so both can be compared row by row.

The default corpus is synthetic: the reference's 639 greyscale JPEGs
(/data/dt_images, benchmarks/distance_transform.cpp:17) are not available
here.  Image i is a seeded blob field (SplitMix64, seed 2000 + i) whose
pixel count equals the i-th distinct ``size`` of the reference's published
results (8.5-11.5 Mpx, 4:3), with a single seed pixel at the centre.  With
``--corpus DIR`` the images are instead the files of DIR in sorted path
order, read with the library's imread (any file the reference's imread
reads: grey JPEG, PNG, PGM, TGA), as the reference lists and loads its
directory (benchmarks/distance_transform.cpp:19-37, :57-59).  The
reference's mask is uninitialised heap memory apart from one pixel
(benchmarks/distance_transform.cpp:66-67); this harness defines it instead.

  python -m distance_transform_benchmarks_amd.gbench --images 4 --out gbench.json
  python -m distance_transform_benchmarks_amd.gbench --corpus /data/dt_images --images 639
  python -m distance_transform_benchmarks_amd.gbench --compare gbench.json REF.json
"""
from __future__ import annotations

import argparse
import datetime
import json
import math
import os
import socket
import statistics
import sys

import numpy as np

LAMBDAS = (0.0, 0.5, 1.0)
FAMILIES = (("DistanceTransformGeos", "raster"), ("DistanceTransformBlocks", "blocks"),
            ("DistanceTransformTask", "task"))

# The first distinct image sizes (pixels) of the reference's published
# results, in the order of its sorted corpus (results/benchmark_results.json,
# "size" counters of DistanceTransformGeos at lambda 0; 116 distinct sizes,
# 8.5-11.5 Mpx, median 8.89 Mpx).
REFERENCE_SIZES = (8505000, 9639000, 9065000, 8823500, 8984500, 8939000, 8897000, 8813000, 8694000, 9023000,
                   8841000, 8806000)


def corpus_shape(n_px: int):
    """4:3 landscape shape with exactly about n_px pixels."""
    h = int(round(math.sqrt(n_px * 3 / 4)))
    w = int(round(n_px / h))
    return w, h


def corpus_image(i: int):
    from . import synth

    w, h = corpus_shape(REFERENCE_SIZES[i % len(REFERENCE_SIZES)])
    rng = synth.SplitMix64(2000 + i)
    img = synth.to_u8(synth.blob_field(rng, w, h))
    mask = np.zeros((h, w), np.uint8)
    mask[h // 2, w // 2] = 1
    return img, mask


def corpus_files(directory: str):
    """The files of a corpus directory in sorted path order."""
    return sorted(os.path.join(directory, f) for f in os.listdir(directory)
                  if os.path.isfile(os.path.join(directory, f)))


def file_image(path: str):
    """A corpus file read with the library's imread, with one seed pixel at
    the centre."""
    from .imageio import imread

    img = imread(path)
    mask = np.zeros(img.shape, np.uint8)
    mask[img.shape[0] // 2, img.shape[1] // 2] = 1
    return img, mask


def run(n_images: int, iterations: int, families, device="cuda:0", corpus: str = ""):
    import torch

    from . import api

    dev = torch.device(device)
    files = corpus_files(corpus) if corpus else []
    if corpus:
        n_images = min(n_images, len(files))
    rows = []
    for fi, (fam, sched) in enumerate(FAMILIES):
        if sched not in families:
            continue
        for i in range(n_images):
            img_np, mask_np = file_image(files[i]) if corpus else corpus_image(i)
            img = torch.from_numpy(img_np).to(dev)
            mask = torch.from_numpy(mask_np).to(dev)
            size = int(img_np.size)
            for j, lam in enumerate(LAMBDAS):
                api.ggdt(img, mask, lam, schedule=sched)  # warm-up (workspace, code objects)
                times = []
                for _ in range(iterations):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record()
                    api.ggdt(img, mask, lam, schedule=sched)  # allocating call, as the reference times it
                    e1.record()
                    e1.synchronize()
                    times.append(e0.elapsed_time(e1))
                real = sum(times) / len(times)
                name = f"{fam}/{i}/{j}/manual_time"
                rows.append({"name": name, "family_index": fi, "per_family_instance_index": i * 3 + j,
                             "run_name": name, "run_type": "iteration", "repetitions": 1, "repetition_index": 0,
                             "threads": 1, "iterations": iterations, "real_time": real, "cpu_time": real,
                             "time_unit": "ms", "bytes_per_second": size / (real * 1e-3), "lambda": lam,
                             "size": float(size),
                             "label": files[i] if corpus else f"synthetic:blob_{i}_{img_np.shape[1]}x{img_np.shape[0]}"})
                print(f"{name}: {real:.3f} ms ({size / real / 1e3:.1f} Mpix/s)", file=sys.stderr, flush=True)
            del img, mask
    ctx = {"date": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
           "host_name": socket.gethostname(), "executable": "python -m distance_transform_benchmarks_amd.gbench",
           "num_cpus": os.cpu_count(), "library_build_type": "release",
           "device": torch.cuda.get_device_name(dev)}
    return {"context": ctx, "benchmarks": rows}


def summarize(doc) -> dict:
    """Median Mpix/s per (family, lambda), the aggregate the reference's
    analysis reports."""
    acc = {}
    for r in doc["benchmarks"]:
        fam = r["name"].split("/")[0]
        acc.setdefault((fam, float(r["lambda"])), []).append(r["bytes_per_second"] / 1e6)
    return {f"{k[0]} lambda={k[1]}": round(statistics.median(v), 1) for k, v in sorted(acc.items())}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="google-benchmark style GGDT benchmark over a synthetic or file corpus")
    ap.add_argument("--images", type=int, default=4)
    ap.add_argument("--corpus", default="", help="directory of images (sorted by path) instead of the synthetic corpus")
    ap.add_argument("--iterations", type=int, default=3)
    ap.add_argument("--families", default="raster,blocks,task")
    ap.add_argument("--out", default="")
    ap.add_argument("--compare", nargs=2, metavar=("OURS", "REFERENCE"),
                    help="print median Mpix/s per family and lambda of two result files")
    a = ap.parse_args(argv)
    if a.compare:
        ours, ref = (summarize(json.load(open(p))) for p in a.compare)
        for k in sorted(set(ours) | set(ref)):
            print(f"{k:<40} ours {ours.get(k, float('nan')):>10} Mpix/s   reference {ref.get(k, float('nan')):>10} Mpix/s")
        return 0
    doc = run(a.images, a.iterations, set(a.families.split(",")), corpus=a.corpus)
    txt = json.dumps(doc, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    else:
        print(txt)
    print(json.dumps(summarize(doc)), file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
