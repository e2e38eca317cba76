"""ctypes wrapper of the CPU oracle (oracle/ggdt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the CPU
baseline, never as the product path.  The product (the HIP library behind
include/dtk.h) does not import or link anything here.

Every function takes and returns contiguous numpy arrays of shape (H, W)
(or (H, W, 3) for RGB8 images).  The restated reference functions are cited
in ggdt_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libggdt_oracle.so")
_lib = None

IMG_U8, IMG_RGB8, IMG_F32 = 0, 1, 2
MASK_U8, MASK_F32 = 0, 1


_REF_ROOT = "/root/reference"
_STB_SRC = os.path.join(_REF_ROOT, "dt", "src", "stb_image_implementation.c")
_STB_PATH = os.path.join(_HERE, "_ref", "libstb_image.so")
_stb = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off), and the
    reference's stb_image into _ref/ when the reference sources are present."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.exists(_STB_SRC):
        subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REF={_REF_ROOT}"], check=True)
    return _LIB_PATH


def stb_available() -> bool:
    """True when the reference's stb_image has been built into oracle/_ref."""
    return os.path.exists(_STB_PATH)


def stb_load(path: str):
    """The reference's imread decoding step (dt/src/imread.cpp:12,
    stbi_load(filename, &w, &h, &n, 0)) through the reference's own
    stb_image built in oracle/_ref.  Returns (pixels, n) with pixels of
    shape (h, w, n), or raises RuntimeError with stb's failure reason."""
    global _stb
    if _stb is None:
        lib = ctypes.CDLL(_STB_PATH)
        P = ctypes.POINTER
        lib.stbi_load.argtypes = [ctypes.c_char_p, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int), ctypes.c_int]
        lib.stbi_load.restype = ctypes.c_void_p
        lib.stbi_image_free.argtypes = [ctypes.c_void_p]
        lib.stbi_failure_reason.restype = ctypes.c_char_p
        _stb = lib
    w, h, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = _stb.stbi_load(str(path).encode(), ctypes.byref(w), ctypes.byref(h), ctypes.byref(n), 0)
    if not p:
        raise RuntimeError(_stb.stbi_failure_reason().decode())
    try:
        size = w.value * h.value * n.value
        px = np.frombuffer((ctypes.c_ubyte * size).from_address(p), np.uint8).copy()
    finally:
        _stb.stbi_image_free(p)
    return px.reshape(h.value, w.value, n.value), n.value


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        vp, fp, i, f, dp = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        lib.ora_init.argtypes = [vp, i, fp, i, i, f]
        lib.ora_raster.argtypes = [vp, i, vp, i, fp, i, i, f, f, i, i]
        lib.ora_raster.restype = i
        lib.ora_raster_omp.argtypes = [vp, i, vp, i, fp, i, i, f, f, i, i, i]
        lib.ora_raster_omp.restype = i
        lib.ora_max_threads.restype = i
        lib.ora_dijkstra.argtypes = [vp, i, vp, i, fp, i, i, f, f, i]
        lib.ora_dijkstra.restype = i
        lib.ora_pred.argtypes = [vp, i, vp, i, fp, vp, i, i, f, f, i]
        lib.ora_backward.argtypes = [fp, fp, vp, fp, vp, vp, vp, i, i, f, f]
        lib.ora_raster_f64.argtypes = [vp, vp, vp, i, i, dp, dp, i]
        lib.ora_raster_f64.restype = i
        lib.ora_fd_grad.argtypes = [fp, fp, fp, i, i, dp, dp, i, dp, dp, vp, i, vp, vp]
        lib.ora_fd_sides.argtypes = [fp, fp, fp, i, i, dp, dp, i, dp, dp, vp, i, vp]
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _img_kind(img: np.ndarray) -> int:
    if img.dtype == np.uint8 and img.ndim == 2:
        return IMG_U8
    if img.dtype == np.uint8 and img.ndim == 3 and img.shape[2] == 3:
        return IMG_RGB8
    if img.dtype == np.float32 and img.ndim == 2:
        return IMG_F32
    raise TypeError(f"unsupported image {img.dtype} {img.shape}")


def _mask_kind(mask: np.ndarray) -> int:
    if mask.dtype == np.uint8:
        return MASK_U8
    if mask.dtype == np.float32:
        return MASK_F32
    raise TypeError(f"unsupported mask {mask.dtype}")


def _prep(img, mask):
    img = np.ascontiguousarray(img)
    mask = np.ascontiguousarray(mask)
    if mask.shape != img.shape[:2]:
        raise ValueError("image and mask shapes differ")
    return img, mask


def init(mask: np.ndarray, v: float) -> np.ndarray:
    mask = np.ascontiguousarray(mask)
    H, W = mask.shape
    D = np.empty((H, W), np.float32)
    _load().ora_init(_ptr(mask), _mask_kind(mask), _ptr(D), W, H, v)
    return D


def raster(img, mask, lam: float, v: float = 1e10, conn8: bool = True, max_iters: int = 0, threads: int = 1):
    """Reference raster GGDT; converged when max_iters <= 0.  Returns (D, iterations)."""
    img, mask = _prep(img, mask)
    H, W = mask.shape
    D = np.empty((H, W), np.float32)
    lib = _load()
    if threads > 1:
        it = lib.ora_raster_omp(_ptr(img), _img_kind(img), _ptr(mask), _mask_kind(mask), _ptr(D), W, H, lam, v,
                                int(conn8), max_iters, threads)
    else:
        it = lib.ora_raster(_ptr(img), _img_kind(img), _ptr(mask), _mask_kind(mask), _ptr(D), W, H, lam, v,
                            int(conn8), max_iters)
    return D, it


def max_threads() -> int:
    return _load().ora_max_threads()


def dijkstra(img, mask, lam: float, v: float = 1e10, conn8: bool = True) -> np.ndarray:
    img, mask = _prep(img, mask)
    H, W = mask.shape
    D = np.empty((H, W), np.float32)
    if _load().ora_dijkstra(_ptr(img), _img_kind(img), _ptr(mask), _mask_kind(mask), _ptr(D), W, H, lam, v,
                            int(conn8)) != 0:
        raise MemoryError("oracle dijkstra allocation failed")
    return D


def pred(img, mask, D, lam: float, v: float, conn8: bool = True) -> np.ndarray:
    img, mask = _prep(img, mask)
    D = np.ascontiguousarray(D, np.float32)
    H, W = mask.shape
    P = np.empty((H, W), np.uint8)
    _load().ora_pred(_ptr(img), _img_kind(img), _ptr(mask), _mask_kind(mask), _ptr(D), _ptr(P), W, H, lam, v,
                     int(conn8))
    return P


def backward(img, D, P, Wt, lam: float, nu: float):
    """Reference gradient of sum(Wt*D): returns (dimg, dmask, G) as float64."""
    img = np.ascontiguousarray(img, np.float32)
    D = np.ascontiguousarray(D, np.float32)
    P = np.ascontiguousarray(P, np.uint8)
    Wt = np.ascontiguousarray(Wt, np.float32)
    H, W = D.shape
    dimg = np.empty((H, W), np.float64)
    dmask = np.empty((H, W), np.float64)
    G = np.empty((H, W), np.float64)
    _load().ora_backward(_ptr(img), _ptr(D), _ptr(P), _ptr(Wt), _ptr(dimg), _ptr(dmask), _ptr(G), W, H, lam, nu)
    return dimg, dmask, G


def raster_f64(img, smask, lam: float, nu: float, conn8: bool = True) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float64)
    smask = np.ascontiguousarray(smask, np.float64)
    H, W = img.shape
    D = np.empty((H, W), np.float64)
    _load().ora_raster_f64(_ptr(img), _ptr(smask), _ptr(D), W, H, lam, nu, int(conn8))
    return D


def fd_grad(img, smask, Wt, lam: float, nu: float, idx, eps: float = 1e-4, conn8: bool = True,
            eps_mask: float = None):
    """FP64 central differences of sum(Wt*D) at flat pixel indices idx
    (image step eps, soft-mask step eps_mask, default eps)."""
    if eps_mask is None:
        eps_mask = eps
    img = np.ascontiguousarray(img, np.float32)
    smask = np.ascontiguousarray(smask, np.float32)
    Wt = np.ascontiguousarray(Wt, np.float32)
    idx = np.ascontiguousarray(idx, np.int64)
    H, W = img.shape
    gi = np.empty(len(idx), np.float64)
    gm = np.empty(len(idx), np.float64)
    _load().ora_fd_grad(_ptr(img), _ptr(smask), _ptr(Wt), W, H, lam, nu, int(conn8), eps, eps_mask, _ptr(idx), len(idx),
                        _ptr(gi), _ptr(gm))
    return gi, gm


def fd_kinkfree(img, smask, Wt, lam: float, nu: float, idx, eps: float = 1e-4, conn8: bool = True,
                eps_mask: float = None, slope_rtol: float = 1e-3):
    """Central differences of sum(Wt*D) at steps eps (image) and eps_mask
    (soft mask), with the pixels where each step stays on one linear piece.
    D is piecewise linear in every input, so when the one-sided slopes
    (L(+) - L) / eps and (L - L(-)) / eps agree to slope_rtol, the central
    difference is the derivative there.  Returns (fd_img, ok_img, fd_mask,
    ok_mask)."""
    if eps_mask is None:
        eps_mask = eps
    img = np.ascontiguousarray(img, np.float32)
    smask = np.ascontiguousarray(smask, np.float32)
    Wt = np.ascontiguousarray(Wt, np.float32)
    idx = np.ascontiguousarray(idx, np.int64)
    H, W = img.shape
    out = np.empty((len(idx), 5), np.float64)
    _load().ora_fd_sides(_ptr(img), _ptr(smask), _ptr(Wt), W, H, lam, nu, int(conn8), eps, eps_mask, _ptr(idx),
                         len(idx), _ptr(out))
    l0 = out[:, 4]
    res = []
    for plus, minus, st in ((out[:, 0], out[:, 1], eps), (out[:, 2], out[:, 3], eps_mask)):
        up, down = (plus - l0) / st, (l0 - minus) / st
        fd = (plus - minus) / (2 * st)
        ok = np.abs(up - down) <= slope_rtol * np.maximum(np.abs(fd), 1e-3 * max(1.0, np.abs(fd).max()))
        res += [fd, ok]
    return tuple(res)


# ---- display of a distance map (numpy restatements) --------------------------
# TEST INFRASTRUCTURE ONLY, like the rest of this module.  Pinned by known
# answers in tests/test_display.py and, for the colour map, by the reference's
# own table (tests/golden/inferno_table.json).

def minmax_f32(D: np.ndarray):
    """dt/include/dt/normalize.hpp:28-43: min starts at FLT_MAX, max at
    numeric_limits<float>::min() (the smallest positive float); a NaN never
    replaces the running value (std::min / std::max with it second)."""
    v = np.asarray(D, np.float32).ravel()
    v = v[~np.isnan(v)]
    mn = np.float32(np.finfo(np.float32).max)
    mx = np.float32(np.finfo(np.float32).tiny)
    if v.size:
        mn = min(mn, v.min())
        mx = max(mx, v.max())
    return np.float32(mn), np.float32(mx)


def normalize_u8(D: np.ndarray) -> np.ndarray:
    """dt/include/dt/normalize.hpp:45-62 with O = uint8: the subtractions in
    float32, the quotient and the scaling by 255 in float64, truncated;
    a NaN quotient (0/0, undefined in the reference) gives 0."""
    D = np.asarray(D, np.float32)
    mn, mx = minmax_f32(D)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        num = (D - mn).astype(np.float64)
        den = np.float64(np.float32(mx - mn))
        v = num / den * 255.0
    out = np.zeros(D.shape, np.uint8)
    ok = ~np.isnan(v)
    out[ok] = np.trunc(v[ok]).astype(np.uint8)
    return out


def inferno_rgb(u8: np.ndarray, table: np.ndarray) -> np.ndarray:
    """dt/src/inferno.cpp:263-274: out = table[v] for a (256, 3) table."""
    return np.asarray(table, np.uint8)[np.asarray(u8, np.uint8)]


def block_priorities(mask: np.ndarray, return_distance: bool = False):
    """dt/src/block_priority.cu:74-155 (dt::compute_block_priorities) on the
    CPU.  32 x 32 tiles (dt/src/utils.cuh:11), gh x gw of them, n = gh * gw:
    - distance 0 for a tile holding a nonzero mask pixel, else uint32 max / 2
      (:84-91, initialize_distance :12-35);
    - the city-block distance on the tile grid: every row forward then
      backward, then every column forward then backward, d = min(d(prev) + 1, d)
      (:37-71, :94-106);
    - cdf[d] = number of tiles at distance <= d (:109-115), level0 = cdf[0];
    - priority = cdf[d] * 64 / (n + 1) in uint32 arithmetic (:118-121);
    - offsets[k] = number of tiles of priority < k (:125-148).
    With no seed tile every distance stays far and every cdf value is n.
    Returns (priorities (gh, gw) uint8, offsets (64,) uint32, level0[, d])."""
    m = np.asarray(mask)
    H, W = m.shape[0], m.shape[1]
    gh, gw = (H + 31) // 32, (W + 31) // 32
    n = gh * gw
    far = 0x7FFFFFFF
    if n == 0:
        out = (np.zeros((gh, gw), np.uint8), np.zeros(64, np.uint32), 0)
        return out + (np.zeros((gh, gw), np.int64),) if return_distance else out
    seed_px = np.zeros((gh * 32, gw * 32), bool)
    seed_px[:H, :W] = m != 0
    seed = seed_px.reshape(gh, 32, gw, 32).any(axis=(1, 3))
    d = np.where(seed, 0, far).astype(np.int64)
    for x in range(1, gw):
        d[:, x] = np.minimum(d[:, x - 1] + 1, d[:, x])
    for x in range(gw - 2, -1, -1):
        d[:, x] = np.minimum(d[:, x + 1] + 1, d[:, x])
    for y in range(1, gh):
        d[y] = np.minimum(d[y - 1] + 1, d[y])
    for y in range(gh - 2, -1, -1):
        d[y] = np.minimum(d[y + 1] + 1, d[y])
    if d.max() >= far:
        cdf_of = np.full(d.shape, n, np.uint64)
        level0 = 0
    else:
        cdf = np.cumsum(np.bincount(d.ravel(), minlength=int(d.max()) + 1))
        cdf_of = cdf[d].astype(np.uint64)
        level0 = int(cdf[0])
    prio = (((cdf_of * np.uint64(64)) % np.uint64(1 << 32)) // np.uint64(n + 1)).astype(np.uint8)
    counts = np.bincount(prio.ravel(), minlength=64)[:64]
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32)
    out = (prio, offsets, level0)
    return out + (d,) if return_distance else out
