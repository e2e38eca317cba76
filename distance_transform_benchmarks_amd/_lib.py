"""ctypes binding of libdtk.so (the C ABI of include/dtk.h).

The library is loaded from this package's ``lib/`` directory.  There is no
fallback: if it is missing, every call raises and names the build command.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIBDTK = os.path.join(LIB_DIR, "libdtk.so")

DTK_OK = 0
DTK_ERR_INVALID_ARGUMENT = 1
DTK_ERR_OUT_OF_MEMORY = 2
DTK_ERR_HIP = 3
DTK_ERR_UNSUPPORTED = 4

DTK_U8, DTK_RGB8, DTK_F32 = 0, 1, 2
DTK_C4, DTK_C8 = 4, 8
SCHED_RASTER, SCHED_BLOCKS, SCHED_TASK = 0, 1, 2
ORDER_DEFAULT, ORDER_FIFO, ORDER_KEY, ORDER_SEED = 0, 1, 2, 3

# Every symbol include/dtk.h declares (checked by the CPU test suite).
EXPORTS = (
    "dtk_params_default", "dtk_last_error", "dtk_version", "dtk_ggdt", "dtk_ggdt_batched",
    "dtk_batch_plan_create", "dtk_batch_plan_launch", "dtk_batch_plan_destroy", "dtk_ggdt_pred",
    "dtk_ggdt_pred_batched", "dtk_ggdt_backward", "dtk_ggdt_backward_batched", "dtk_tile_count", "dtk_malloc_pitch", "dtk_free", "dtk_host_to_device",
    "dtk_device_to_host", "dtk_copy2d", "dtk_selftest", "dtk_device_cu_count", "dtk_profile_begin",
    "dtk_profile_end", "dtk_batch_autotune", "dtk_tile_count_for", "dtk_minmax_f32", "dtk_normalize_u8",
    "dtk_inferno", "dtk_colorize", "dtk_block_priorities", "dtk_sync_errors", "dtk_tile_size_for",
    "dtk_tile_image", "dtk_ggdt_noalloc",
)


class DtkImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("pitch", ctypes.c_longlong), ("dtype", ctypes.c_int)]


class DtkParams(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_float), ("v", ctypes.c_float), ("connectivity", ctypes.c_int),
                ("schedule", ctypes.c_int), ("max_iterations", ctypes.c_int), ("tile_rows", ctypes.c_int),
                ("max_sweeps", ctypes.c_int), ("waves_per_cu", ctypes.c_int), ("cost_plane", ctypes.c_int),
                ("tile_waves", ctypes.c_int), ("queue_order", ctypes.c_int),
                ("round_per_blocks", ctypes.c_void_p)]


class DtkStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_int), ("visits", ctypes.c_uint32), ("sweeps", ctypes.c_uint32),
                ("tiles", ctypes.c_int), ("tile_rows", ctypes.c_int), ("workers", ctypes.c_int),
                ("cycles_wait", ctypes.c_uint32), ("cycles_visit", ctypes.c_uint32), ("cycles_pub", ctypes.c_uint32),
                ("cycles_setup", ctypes.c_uint32), ("cycles_sweep", ctypes.c_uint32), ("cycles_store", ctypes.c_uint32),
                ("cycles_release", ctypes.c_uint32), ("cycles_acquire", ctypes.c_uint32),
                ("drains", ctypes.c_uint32), ("moved", ctypes.c_uint32), ("deferred", ctypes.c_uint32),
                ("cycles_drain", ctypes.c_uint32), ("kept", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DtkTuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_int), ("from_cache", ctypes.c_int), ("best_ms", ctypes.c_float),
                ("default_ms", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DtkError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"dtk status {status}: {msg}")
        self.status = status


_lib = None


def load() -> ctypes.CDLL:
    """Load libdtk.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBDTK):
        raise RuntimeError(f"{LIBDTK} is missing: build it with `python -m distance_transform_benchmarks_amd.build`")
    lib = ctypes.CDLL(LIBDTK)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.dtk_params_default.argtypes = [P(DtkParams)]
    lib.dtk_params_default.restype = None
    lib.dtk_last_error.restype = ctypes.c_char_p
    lib.dtk_version.restype = ctypes.c_char_p
    lib.dtk_ggdt.argtypes = [P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp, P(DtkStats)]
    lib.dtk_ggdt_batched.argtypes = [ctypes.c_int, P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp,
                                     P(DtkStats)]
    lib.dtk_batch_plan_create.argtypes = [ctypes.c_int, P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp,
                                          P(vp)]
    lib.dtk_batch_plan_launch.argtypes = [vp, vp]
    lib.dtk_batch_plan_destroy.argtypes = [vp]
    lib.dtk_batch_autotune.argtypes = [ctypes.c_int, P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp,
                                       ctypes.c_int, P(DtkParams), P(DtkTuneReport)]
    lib.dtk_ggdt_pred.argtypes = [P(DtkImage), P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp]
    lib.dtk_ggdt_pred_batched.argtypes = [ctypes.c_int, P(DtkImage), P(DtkImage), P(DtkImage), P(DtkImage),
                                          P(DtkParams), vp]
    lib.dtk_ggdt_backward.argtypes = [P(DtkImage), P(DtkImage), P(DtkImage), P(DtkImage), P(DtkImage),
                                      P(DtkImage), P(DtkImage), P(DtkParams), vp]
    lib.dtk_ggdt_backward_batched.argtypes = [ctypes.c_int, P(DtkImage), P(DtkImage), P(DtkImage), P(DtkImage),
                                              P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp]
    lib.dtk_tile_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.dtk_tile_count.restype = ctypes.c_int
    lib.dtk_tile_count_for.argtypes = [ctypes.c_int, ctypes.c_int, P(DtkParams)]
    lib.dtk_tile_count_for.restype = ctypes.c_int
    lib.dtk_malloc_pitch.argtypes = [P(vp), P(ctypes.c_longlong), ctypes.c_longlong, ctypes.c_int]
    lib.dtk_free.argtypes = [vp]
    lib.dtk_host_to_device.argtypes = [P(DtkImage), P(DtkImage)]
    lib.dtk_device_to_host.argtypes = [P(DtkImage), P(DtkImage)]
    lib.dtk_copy2d.argtypes = [vp, ctypes.c_longlong, vp, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                               ctypes.c_int]
    lib.dtk_selftest.argtypes = [P(ctypes.c_int)]
    lib.dtk_device_cu_count.restype = ctypes.c_int
    lib.dtk_profile_begin.argtypes = []
    lib.dtk_profile_end.argtypes = [P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int)]
    lib.dtk_minmax_f32.argtypes = [P(DtkImage), P(ctypes.c_float), vp]
    lib.dtk_normalize_u8.argtypes = [P(DtkImage), P(DtkImage), vp]
    lib.dtk_inferno.argtypes = [P(DtkImage), P(DtkImage), vp]
    lib.dtk_colorize.argtypes = [P(DtkImage), P(DtkImage), vp]
    lib.dtk_block_priorities.argtypes = [P(DtkImage), vp, vp, P(ctypes.c_uint32), vp]
    lib.dtk_sync_errors.argtypes = [vp]
    lib.dtk_ggdt_noalloc.argtypes = [P(DtkImage), P(DtkImage), P(DtkImage), P(DtkParams), vp, P(DtkStats)]
    lib.dtk_tile_size_for.argtypes = [P(DtkParams), P(ctypes.c_int), P(ctypes.c_int)]
    lib.dtk_tile_size_for.restype = ctypes.c_int
    lib.dtk_tile_image.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   P(DtkImage), vp]
    for name in ("dtk_ggdt", "dtk_ggdt_batched", "dtk_batch_plan_create", "dtk_batch_plan_launch",
                 "dtk_batch_plan_destroy", "dtk_batch_autotune", "dtk_ggdt_pred", "dtk_ggdt_pred_batched", "dtk_ggdt_backward", "dtk_ggdt_backward_batched", "dtk_malloc_pitch", "dtk_free",
                 "dtk_host_to_device", "dtk_device_to_host", "dtk_copy2d", "dtk_selftest", "dtk_profile_begin",
                 "dtk_profile_end", "dtk_sync_errors", "dtk_block_priorities", "dtk_minmax_f32", "dtk_normalize_u8",
                 "dtk_inferno", "dtk_colorize", "dtk_tile_image", "dtk_ggdt_noalloc"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != DTK_OK:
        msg = load().dtk_last_error().decode(errors="replace")
        if status == DTK_ERR_INVALID_ARGUMENT:
            raise ValueError(msg)
        raise DtkError(status, msg)


def default_params() -> DtkParams:
    p = DtkParams()
    load().dtk_params_default(ctypes.byref(p))
    return p
