"""Image files for the Python side: the reference's ``dt::imread<uint8_t>``
and ``dt::imsave`` (dt/src/imread.cpp:9-29, dt/src/imsave.cpp:9-30) over
the C ABI of include/dt_io.h in lib/libdt.so.

``imread`` accepts what the reference's imread accepts: any file its
stb_image decodes to exactly one 8-bit component (grey PNG, JPEG, PGM, TGA).
A file that cannot be read raises ``RuntimeError("Error while loading
<file>...")``, and an image with another component count raises
``ValueError("Input image must have 1 component (grayscale) of 8 bits")``,
the reference's std::runtime_error / std::invalid_argument.  ``imsave``
writes 8-bit grey (H, W) or RGB (H, W, 3) arrays as PNG and refuses other
file names with ``ValueError("Only PNG output is supported")``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import LIB_DIR

LIBDT = os.path.join(LIB_DIR, "libdt.so")
_IO_OK, _IO_ERR_READ, _IO_ERR_FORMAT = 0, 1, 2
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBDT):
            raise RuntimeError(f"{LIBDT} is missing: build the package first (python -m "
                               "distance_transform_benchmarks_amd.build)")
        lib = ctypes.CDLL(LIBDT)
        P = ctypes.POINTER
        lib.dt_png_info.argtypes = [ctypes.c_char_p, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
        lib.dt_png_read_u8.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_longlong]
        lib.dt_png_write.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_longlong, ctypes.c_int]
        lib.dt_io_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _raise(lib, status: int):
    msg = lib.dt_io_last_error().decode()
    if status == _IO_ERR_FORMAT:
        raise ValueError(msg)
    raise RuntimeError(msg)


def image_info(path) -> tuple:
    """(width, height, components) of an image file, components counted as
    the reference's decoder counts them (palette PNG 3 or 4, JPEG 1 or 3,
    BMP/GIF/PSD 3 or 4, ...)."""
    lib = _load()
    w, h, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    r = lib.dt_png_info(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c))
    if r != _IO_OK:
        _raise(lib, r)
    return w.value, h.value, c.value


def imread(path) -> np.ndarray:
    """dt::imread<uint8_t>: a one-component 8-bit image as an (H, W) uint8 array."""
    lib = _load()
    w, h, c = image_info(path)
    if c != 1:
        raise ValueError("Input image must have 1 component (grayscale) of 8 bits")
    out = np.empty((h, w), np.uint8)
    r = lib.dt_png_read_u8(os.fsencode(path), out.ctypes.data, w, h, w)
    if r != _IO_OK:
        _raise(lib, r)
    return out


def imsave(path, image: np.ndarray) -> None:
    """dt::imsave: an (H, W) grey or (H, W, 3) RGB uint8 array as PNG."""
    lib = _load()
    a = np.ascontiguousarray(image, dtype=np.uint8)
    if a.ndim == 2:
        channels = 1
    elif a.ndim == 3 and a.shape[2] == 3:
        channels = 3
    else:
        raise ValueError(f"imsave: expected an (H, W) or (H, W, 3) uint8 array, got shape {a.shape}")
    r = lib.dt_png_write(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0], a.shape[1] * channels, channels)
    if r != _IO_OK:
        _raise(lib, r)
