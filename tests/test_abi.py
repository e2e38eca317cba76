"""The C-ABI boundary on the CPU: the library loads, exports what the header
declares, lays out its structs as the header does, and rejects invalid
arguments before touching a device.  No compute call is made here."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dtk.h")
PKG = os.path.join(ROOT, "distance_transform_benchmarks_amd")


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(dtk_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from distance_transform_benchmarks_amd import _lib

    if not os.path.exists(_lib.LIBDTK):
        pytest.fail(f"{_lib.LIBDTK} missing: run __graft_entry__.build() first")
    return _lib.load()


def test_header_declares_the_binding_exports():
    from distance_transform_benchmarks_amd import _lib

    declared = _header_functions()
    assert len(declared) >= 20
    assert sorted(_lib.EXPORTS) == declared


def test_library_exports_every_declared_symbol(lib):
    from distance_transform_benchmarks_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBDTK], capture_output=True, text=True, check=True)
    exported = set(line.split()[-1] for line in out.stdout.splitlines() if line.strip())
    missing = [s for s in _header_functions() if s not in exported]
    assert not missing, missing
    for s in _header_functions():
        assert hasattr(lib, s)


def test_cpp_api_library_exports_reference_entry_points():
    so = os.path.join(PKG, "lib", "libdt.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    for name in ("dt::generalised_distance_transform(", "dt::generalised_distance_transform_blocks(",
                 "dt::generalised_distance_transform_task(", "dt::host_to_device(", "dt::device_to_host("):
        assert name in out, name


def test_cpp_api_library_exports_display_and_io():
    """The C++ API of the CLI's helpers (imread / imsave / normalize /
    inferno / invert_mask) and the C ABI of include/dt_io.h."""
    so = os.path.join(PKG, "lib", "libdt.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    for name in ("dt::imread<unsigned char>(", "dt::impl::imsave<unsigned char>(", "dt::impl::imsave<dt::rgb8>(",
                 "dt::normalize<unsigned char>(", "dt::minmax(", "dt::inferno(", "dt::invert_mask(",
                 "dt_png_info", "dt_png_read_u8", "dt_png_write", "dt_io_last_error"):
        assert name in out, name
    header = open(os.path.join(ROOT, "include", "dt_io.h")).read()
    for fn in re.findall(r"^\w[\w\s\*]*?\b(dt_\w+)\(", header, re.M):
        assert fn in out, fn
    for tool in ("dt_distance_transform", "dt_empiric_check"):
        assert os.access(os.path.join(PKG, "lib", tool), os.X_OK), tool


def test_struct_layout_matches_header(tmp_path):
    """sizeof/offsetof of the ABI structs as a C compiler sees include/dtk.h
    equal the ctypes mirrors the Python binding uses."""
    from distance_transform_benchmarks_amd._lib import DtkImage, DtkParams, DtkStats

    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "dtk.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(dtk_image), offsetof(dtk_image, pitch), offsetof(dtk_image, dtype),"
        "sizeof(dtk_params), offsetof(dtk_params, round_per_blocks), offsetof(dtk_params, waves_per_cu),"
        "sizeof(dtk_stats), offsetof(dtk_stats, cycles_acquire)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(DtkImage), DtkImage.pitch.offset, DtkImage.dtype.offset, ctypes.sizeof(DtkParams),
            DtkParams.round_per_blocks.offset, DtkParams.waves_per_cu.offset, ctypes.sizeof(DtkStats),
            DtkStats.cycles_acquire.offset]
    assert got == want


def test_params_default_and_version(lib):
    from distance_transform_benchmarks_amd import _lib

    p = _lib.default_params()
    assert p.lambda_ == 0.0 and p.v == pytest.approx(1e10) and p.connectivity == 8
    assert p.schedule == _lib.SCHED_TASK and p.max_iterations == 0 and p.tile_rows == 0
    assert b"gfx950" in lib.dtk_version()


@pytest.mark.parametrize("W,H,rows", [(1, 1, 0), (62, 30, 32), (63, 31, 32), (4096, 4096, 0), (4096, 4096, 16),
                                      (124, 14, 16), (125, 15, 16)])
def test_tile_count_closed_form(lib, W, H, rows):
    rh = rows or 32
    assert lib.dtk_tile_count(W, H, rows) == -(-W // 62) * -(-H // (rh - 2))


@pytest.mark.parametrize("sched,lam,rows,waves,owned_rows", [
    (2, 0.5, 0, 0, 56), (2, 0.0, 0, 0, 120), (2, 0.5, 32, 0, 60), (2, 0.0, 32, 0, 120), (2, 0.0, 16, 0, 56),
    (2, 0.5, 16, 2, 28), (2, 0.0, 32, 1, 30), (1, 0.0, 0, 0, 120), (1, 0.5, 0, 0, 56), (1, 0.5, 32, 1, 30),
    (1, 0.5, 16, 2, 28), (0, 0.5, 16, 0, 14)])
def test_tile_count_for_schedules(lib, sched, lam, rows, waves, owned_rows):
    """Default tilings: task 16-row waves stacked four high (lambda > 0) or
    32-row waves stacked four high (lambda 0); blocks single 32-row waves."""
    from distance_transform_benchmarks_amd import _lib

    p = _lib.default_params()
    p.schedule, p.lambda_, p.tile_rows, p.tile_waves = sched, lam, rows, waves
    W, H = 4096, 4096
    assert lib.dtk_tile_count_for(W, H, ctypes.byref(p)) == -(-W // 62) * -(-H // owned_rows)


def test_tile_count_rejects_bad_rows(lib):
    assert lib.dtk_tile_count(10, 10, 8) == 0
    assert lib.dtk_tile_count(0, 10, 0) == 0


def _img(w, h, dtype, data=0x1000, pitch=None):
    from distance_transform_benchmarks_amd._lib import DtkImage

    es = {0: 1, 1: 3, 2: 4}.get(dtype, 1)
    return DtkImage(data, w, h, pitch if pitch is not None else w * es, dtype)


@pytest.mark.parametrize("case,msg", [
    ("lambda", "lambda"), ("nu", "v must"), ("conn", "connectivity"), ("sched", "schedule"),
    ("rows", "tile_rows"), ("fixed_task", "fixed iteration"), ("size", "sizes differ"),
    ("dtype", "unsupported dtype"), ("pitch", "pitch"), ("null", "NULL"), ("align", "aligned"),
])
def test_invalid_arguments_rejected_on_host(lib, case, msg):
    """Every check happens before the first HIP call, so these run without a GPU."""
    from distance_transform_benchmarks_amd import _lib

    p = _lib.default_params()
    img, mask, dist = _img(8, 8, 0), _img(8, 8, 0), _img(8, 8, 2)
    if case == "lambda":
        p.lambda_ = 1.5
    elif case == "nu":
        p.v = float("inf")
    elif case == "conn":
        p.connectivity = 6
    elif case == "sched":
        p.schedule = 7
    elif case == "rows":
        p.tile_rows = 64
    elif case == "fixed_task":
        p.max_iterations = 3
    elif case == "size":
        mask = _img(8, 9, 0)
    elif case == "dtype":
        mask = _img(8, 8, 1)
    elif case == "pitch":
        img = _img(8, 8, 0, pitch=4)
    elif case == "null":
        img = _img(8, 8, 0, data=0)
    elif case == "align":
        dist = _img(8, 8, 2, data=0x1002)
    st = lib.dtk_ggdt(ctypes.byref(img), ctypes.byref(mask), ctypes.byref(dist), ctypes.byref(p), None, None)
    assert st == _lib.DTK_ERR_INVALID_ARGUMENT
    assert msg in lib.dtk_last_error().decode()
    with pytest.raises(ValueError):
        _lib.check(st)


def test_priority_and_batched_backward_reject_bad_arguments_on_host(lib):
    """dtk_block_priorities and dtk_ggdt_backward_batched check their
    arguments before the first HIP call."""
    from distance_transform_benchmarks_amd import _lib

    lv = ctypes.c_uint32(7)
    rgb = _img(8, 8, 1)
    assert lib.dtk_block_priorities(ctypes.byref(rgb), None, None, ctypes.byref(lv), None) == _lib.DTK_ERR_INVALID_ARGUMENT
    assert "unsupported dtype" in lib.dtk_last_error().decode()
    m = _img(8, 8, 0)
    assert lib.dtk_block_priorities(ctypes.byref(m), None, None, None, None) == _lib.DTK_ERR_INVALID_ARGUMENT
    assert "level0" in lib.dtk_last_error().decode()
    assert lib.dtk_block_priorities(ctypes.byref(m), None, None, ctypes.byref(lv), None) == _lib.DTK_ERR_INVALID_ARGUMENT
    assert "NULL" in lib.dtk_last_error().decode()
    empty = _img(0, 5, 0)
    assert lib.dtk_block_priorities(ctypes.byref(empty), None, None, ctypes.byref(lv), None) == _lib.DTK_OK
    assert lv.value == 0

    p = _lib.default_params()
    p.lambda_ = 1.0

    def arr(*ims):
        return (_lib.DtkImage * len(ims))(*ims)

    f32 = lambda w=8, h=8: _img(w, h, 2)  # noqa: E731
    u8 = lambda w=8, h=8: _img(w, h, 0)  # noqa: E731
    assert lib.dtk_ggdt_backward_batched(-1, None, None, None, None, None, None, None, ctypes.byref(p), None) == \
        _lib.DTK_ERR_INVALID_ARGUMENT
    assert lib.dtk_ggdt_backward_batched(1, None, None, None, None, None, None, None, ctypes.byref(p), None) == \
        _lib.DTK_ERR_INVALID_ARGUMENT
    assert "NULL" in lib.dtk_last_error().decode()
    assert lib.dtk_ggdt_backward_batched(0, None, None, None, None, None, None, None, ctypes.byref(p), None) == \
        _lib.DTK_OK
    cases = [
        (arr(u8(), u8()), arr(f32(), f32()), "float32 image", _lib.DTK_ERR_UNSUPPORTED),
        (arr(f32(), f32()), arr(f32(), u8()), "share a dtype", _lib.DTK_ERR_INVALID_ARGUMENT),
    ]
    for imgs, masks, msg, code in cases:
        st = lib.dtk_ggdt_backward_batched(2, imgs, masks, arr(f32(), f32()), arr(u8(), u8()), arr(f32(), f32()),
                                           arr(f32(), f32()), None, ctypes.byref(p), None)
        assert st == code and msg in lib.dtk_last_error().decode(), (msg, lib.dtk_last_error())
    st = lib.dtk_ggdt_backward_batched(2, arr(f32(), f32()), arr(f32(), f32()), arr(f32(), f32()), arr(u8(), u8()),
                                       arr(f32(), f32(9, 8)), arr(f32(), f32()), None, ctypes.byref(p), None)
    assert st == _lib.DTK_ERR_INVALID_ARGUMENT and "sizes differ" in lib.dtk_last_error().decode()


def test_cpp_suite_cpu_part():
    """The C++ API test runner (mirrors the reference's gtest cases) without --gpu."""
    exe = os.path.join(PKG, "lib", "dt_tests")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout


def test_gbench_corpus_and_summary():
    """The benchmark harness's synthetic corpus matches the reference's image
    sizes and its summary is the median Mpix/s per family and lambda."""
    from distance_transform_benchmarks_amd import gbench

    for n in gbench.REFERENCE_SIZES:
        w, h = gbench.corpus_shape(n)
        assert abs(w * h - n) <= w and 1.3 < w / h < 1.37
    doc = {"benchmarks": [
        {"name": "DistanceTransformTask/0/1/manual_time", "lambda": 0.5, "bytes_per_second": 3e8},
        {"name": "DistanceTransformTask/1/1/manual_time", "lambda": 0.5, "bytes_per_second": 5e8},
        {"name": "DistanceTransformTask/2/1/manual_time", "lambda": 0.5, "bytes_per_second": 4e8},
        {"name": "DistanceTransformGeos/0/0/manual_time", "lambda": 0.0, "bytes_per_second": 1e7}]}
    assert gbench.summarize(doc) == {"DistanceTransformGeos lambda=0.0": 10.0, "DistanceTransformTask lambda=0.5": 400.0}


def test_batched_predecessors_reject_bad_arguments_on_host(lib):
    """dtk_ggdt_pred_batched checks its arguments before the first HIP call."""
    from distance_transform_benchmarks_amd import _lib

    p = _lib.DtkParams()
    lib.dtk_params_default(ctypes.byref(p))
    arr = lambda *ims: (_lib.DtkImage * len(ims))(*ims)  # noqa: E731
    f32 = lambda w=8, h=8: _img(w, h, 2)  # noqa: E731
    u8 = lambda w=8, h=8: _img(w, h, 0)  # noqa: E731
    assert lib.dtk_ggdt_pred_batched(-1, None, None, None, None, ctypes.byref(p), None) == _lib.DTK_ERR_INVALID_ARGUMENT
    assert lib.dtk_ggdt_pred_batched(1, None, None, None, None, ctypes.byref(p), None) == _lib.DTK_ERR_INVALID_ARGUMENT
    assert "NULL" in lib.dtk_last_error().decode()
    assert lib.dtk_ggdt_pred_batched(0, None, None, None, None, ctypes.byref(p), None) == _lib.DTK_OK
    cases = [
        (arr(u8(), f32()), arr(u8(), u8()), arr(f32(), f32()), arr(u8(), u8()), "share image and mask dtypes"),
        (arr(u8(), u8()), arr(u8(), u8()), arr(f32(), f32()), arr(u8(), u8(9)), "pred size differs"),
        (arr(u8(), u8()), arr(u8(), u8()), arr(f32(), f32()), arr(u8(), f32()), "pred"),
    ]
    for imgs, masks, dists, preds, msg in cases:
        assert lib.dtk_ggdt_pred_batched(2, imgs, masks, dists, preds, ctypes.byref(p), None) == \
            _lib.DTK_ERR_INVALID_ARGUMENT, msg
        assert msg in lib.dtk_last_error().decode(), (msg, lib.dtk_last_error())
