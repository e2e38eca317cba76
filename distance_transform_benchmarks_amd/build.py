"""Build recipe of the native libraries (in-tree, gfx950).

  lib/libdtk.so   HIP kernels + the C ABI of include/dtk.h
  lib/libdt.so    C++ API of include/dt/ (namespace dt), on top of libdtk
  lib/dt_tests    self-contained C++ test runner of the C++ API
  lib/dtk_queue_check  CPU model check of the work-queue protocol

All HIP code is compiled for gfx950 with -ffp-contract=off (the relaxation
must not be contracted into FMAs: bit-exactness with the CPU fixed point)
and without packed FP32 instructions.

Usage: python -m distance_transform_benchmarks_amd.build  [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DTK_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

DTK_SOURCES = ["dtk_task.hip", "dtk_raster.hip", "dtk_adjoint.hip", "dtk_color.hip", "dtk_priority.hip", "dtk_host.cpp"]
DT_SOURCES = ["dt_api.cpp", "dt_color.cpp", "dt_png.cpp", "dt_jpeg.cpp", "dt_image_formats.cpp"]
# Command-line tools (the reference's bin/distance_transform.cpp,
# bin/empiric_check.cpp, bin/rounds_per_blocks.cu and bin/priority_block.cu),
# linked against libdt.
TOOLS = {"dt_distance_transform": "dt_cli_distance_transform.cpp", "dt_empiric_check": "dt_cli_empiric_check.cpp",
         "dt_rounds_per_blocks": "dt_cli_rounds_per_blocks.cpp", "dt_priority_block": "dt_cli_priority_block.cpp"}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build(jobs: int = 8, verbose: bool = True) -> dict:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = DTK_SOURCES + DT_SOURCES
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    libdtk = os.path.join(LIB, "libdtk.so")
    tmp = libdtk + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[objs[s] for s in DTK_SOURCES], "-o", tmp])
    os.replace(tmp, libdtk)
    libdt = os.path.join(LIB, "libdt.so")
    tmp = libdt + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[objs[s] for s in DT_SOURCES], "-o", tmp,
          f"-L{LIB}", "-ldtk", "-lz", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, libdt)
    tools = {}
    for name, src in TOOLS.items():
        exe = os.path.join(LIB, name)
        tmp = exe + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-x", "hip", os.path.join(CSRC, src), "-o", tmp,
              f"-L{LIB}", "-ldt", "-ldtk", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, exe)
        tools[name] = exe
    # CPU model check of the work-queue protocol (host threads only).
    qcheck = os.path.join(LIB, "dtk_queue_check")
    tmp = qcheck + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-x", "hip", os.path.join(CSRC, "dtk_queue_check.cpp"), "-o", tmp,
          "-pthread"])
    os.replace(tmp, qcheck)
    tests = os.path.join(LIB, "dt_tests")
    tmp = tests + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-x", "hip", os.path.join(CSRC, "dt_tests.cpp"), "-o", tmp,
          f"-L{LIB}", "-ldt", "-ldtk", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, tests)
    if verbose:
        print(f"built {libdtk}, {libdt}, {tests}, " + ", ".join(tools.values()))
    return {"libdtk": libdtk, "libdt": libdt, "dt_tests": tests, "dtk_queue_check": qcheck, **tools}


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    shutil.rmtree(LIB, ignore_errors=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args(argv)
    if a.clean:
        clean()
    build(a.jobs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
