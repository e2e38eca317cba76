"""Python mirror of the reference GGDT API over torch device tensors.

Names and argument meaning follow dt/include/dt/generalised_distance_transform.hpp:7-27:

  generalised_distance_transform(img, mask, lambda_, v=1e10)         raster schedule
  generalised_distance_transform_blocks(img, mask, lambda_, v=1e10)  host-driven tile rounds
  generalised_distance_transform_task(img, mask, lambda_, v=1e10)    persistent tile queue

``img`` is a CUDA (HIP) tensor: uint8 (H, W) grey, uint8 (H, W, 3) RGB8 or
float32 (H, W).  ``mask`` is uint8 (H, W) (mask != 0 marks a seed, D0 =
v * (mask == 0), dt/src/utils.cu:12) or float32 (H, W) soft mask s in [0, 1]
(D0 = v * (1 - s)).  The result is a float32 (H, W) tensor; all three
schedules give the same bit-exact converged map.  Like the reference the
functions validate ``0 <= lambda <= 1``; invalid arguments raise ValueError,
device failures raise DtkError.

Extension keywords (no reference counterpart): ``connectivity`` (4 or 8),
``max_iterations`` (raster schedule, fixed iteration count), ``tile_rows``,
``max_sweeps``, ``waves_per_cu`` (tile-solver tuning), ``out`` (write into
an existing float32 tensor) and ``stats`` (return call statistics).

Every call runs on torch's current stream; the task schedule does not
synchronise with the host.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _lib
from ._lib import DtkImage, DtkParams, DtkStats, check

__all__ = [
    "generalised_distance_transform", "generalised_distance_transform_blocks",
    "generalised_distance_transform_task", "ggdt", "ggdt_batched", "BatchPlan", "predecessors",
    "backward", "selftest", "tile_count", "round_per_blocks_buffer", "tile_grid", "visits_buffer",
    "minmax", "normalize_u8", "inferno", "colorize", "synchronize", "tile_size", "tile_image", "heatmap",
]


def _stream_handle(device) -> ctypes.c_void_p:
    s = torch.cuda.current_stream(device)
    return ctypes.c_void_p(s.cuda_stream)


def _as_image(t: torch.Tensor, kind: str) -> DtkImage:
    """Describe a 2-D pitched device tensor (rows may be padded)."""
    if not t.is_cuda:
        raise ValueError(f"{kind} must be a device tensor")
    if t.dtype == torch.uint8 and t.dim() == 3 and t.shape[2] == 3:
        if t.stride(2) != 1 or t.stride(1) != 3:
            raise ValueError(f"{kind}: RGB8 rows must be packed r,g,b")
        return DtkImage(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), _lib.DTK_RGB8)
    if t.dim() != 2:
        raise ValueError(f"{kind} must be (H, W) (or (H, W, 3) uint8)")
    if t.stride(1) != 1:
        raise ValueError(f"{kind}: rows must be contiguous")
    if t.dtype == torch.uint8:
        return DtkImage(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0), _lib.DTK_U8)
    if t.dtype == torch.float32:
        return DtkImage(t.data_ptr(), t.shape[1], t.shape[0], t.stride(0) * 4, _lib.DTK_F32)
    raise ValueError(f"{kind}: unsupported dtype {t.dtype}")


_SCHED = {"raster": _lib.SCHED_RASTER, "blocks": _lib.SCHED_BLOCKS, "task": _lib.SCHED_TASK}


_ORDER = {None: _lib.ORDER_DEFAULT, "default": _lib.ORDER_DEFAULT, "fifo": _lib.ORDER_FIFO, "key": _lib.ORDER_KEY,
          "seed": _lib.ORDER_SEED}


def _params(lambda_, v, connectivity, schedule, max_iterations, tile_rows, max_sweeps, waves_per_cu,
            round_per_blocks=None, cost_plane=False, tile_waves=0, queue_order=None) -> DtkParams:
    p = _lib.default_params()
    p.lambda_ = float(lambda_)
    p.v = float(v)
    if connectivity not in (4, 8):
        raise ValueError("connectivity must be 4 or 8")
    p.connectivity = connectivity
    p.schedule = _SCHED[schedule]
    p.max_iterations = int(max_iterations)
    p.tile_rows = int(tile_rows)
    p.max_sweeps = int(max_sweeps)
    p.waves_per_cu = int(waves_per_cu)
    p.cost_plane = 1 if cost_plane else 0
    p.tile_waves = int(tile_waves)
    if queue_order not in _ORDER:
        raise ValueError(f"queue_order must be one of {sorted(k for k in _ORDER if k)}")
    p.queue_order = _ORDER[queue_order]
    if round_per_blocks is not None:
        if round_per_blocks.dtype != torch.int32 or not round_per_blocks.is_cuda:
            raise ValueError("round_per_blocks must be an int32 device tensor")
        p.round_per_blocks = round_per_blocks.data_ptr()
    return p


def ggdt(img: torch.Tensor, mask: torch.Tensor, lambda_: float, v: float = 1e10, *, connectivity: int = 8,
         schedule: str = "task", max_iterations: int = 0, tile_rows: int = 0, max_sweeps: int = 0,
         waves_per_cu: int = 0, out: Optional[torch.Tensor] = None, stats: bool = False,
         round_per_blocks: Optional[torch.Tensor] = None, cost_plane: bool = False, tile_waves: int = 0,
         queue_order: Optional[str] = None):
    """GGDT of one image (any schedule).  Returns D, or (D, stats dict).

    queue_order (task schedule): "fifo" (the default for one image), "key"
    or "seed"; the order in which dirty tiles are solved, which never changes
    the result."""
    H, W = mask.shape[0], mask.shape[1]
    if img.shape[0] != H or img.shape[1] != W:
        raise ValueError("image and mask sizes differ")
    if out is None:
        out = torch.empty((H, W), dtype=torch.float32, device=img.device)
    p = _params(lambda_, v, connectivity, schedule, max_iterations, tile_rows, max_sweeps, waves_per_cu,
                round_per_blocks, cost_plane, tile_waves, queue_order)
    if round_per_blocks is not None:
        need = _lib.load().dtk_tile_count_for(W, H, ctypes.byref(p))
        if not round_per_blocks.is_contiguous() or round_per_blocks.numel() < need:
            raise ValueError(f"round_per_blocks needs {need} contiguous int32 counters for this call")
    di, dm, dd = _as_image(img, "image"), _as_image(mask, "mask"), _as_image(out, "dist")
    st = DtkStats()
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_ggdt(ctypes.byref(di), ctypes.byref(dm), ctypes.byref(dd), ctypes.byref(p),
                                   _stream_handle(img.device), ctypes.byref(st) if stats else None))
    return (out, st.as_dict()) if stats else out


def generalised_distance_transform(img, mask, lambda_, v=1e10, **kw):
    """Reference raster schedule (dt/src/generalised_distance_transform.cu:101-148)."""
    return ggdt(img, mask, lambda_, v, schedule="raster", **kw)


def generalised_distance_transform_blocks(img, mask, lambda_, v=1e10, round_per_blocks=None, **kw):
    """Reference blocks schedule (dt/src/generalised_distance_transform_blocks.cu:114-184)."""
    return ggdt(img, mask, lambda_, v, schedule="blocks", round_per_blocks=round_per_blocks, **kw)


def generalised_distance_transform_task(img, mask, lambda_, v=1e10, **kw):
    """Reference task schedule (dt/src/generalised_distance_transform_task.cu:152-207)."""
    return ggdt(img, mask, lambda_, v, schedule="task", **kw)


def tile_count(width: int, height: int, tile_rows: int = 0) -> int:
    return _lib.load().dtk_tile_count(width, height, tile_rows)


def round_per_blocks_buffer(width: int, height: int, device="cuda") -> torch.Tensor:
    """Zeroed int32 visit counters for the blocks schedule's instrumentation:
    dtk_tile_count(W, H, 0) entries, enough for any default tiling (the
    first tiles_y * tiles_x of them are used; ``visits_buffer`` gives the
    exact (tiles_y, tiles_x) grid of a call)."""
    return torch.zeros(tile_count(width, height, 0), dtype=torch.int32, device=device)


def tile_grid(width: int, height: int, lambda_: float, *, schedule: str = "task", tile_rows: int = 0,
              tile_waves: int = 0, max_sweeps: int = 0):
    """(tiles_y, tiles_x) of a call's tiling: the shape of its per-tile visit
    counts (``round_per_blocks``).  The task schedule's default tiling
    depends on lambda_ and max_sweeps (lambda 0 without a sweep cap uses
    taller tiles), so pass the call's values."""
    p = _params(lambda_, 1e10, 8, schedule, 0, tile_rows, max_sweeps, 0, tile_waves=tile_waves)
    n = _lib.load().dtk_tile_count_for(width, height, ctypes.byref(p))
    tx = -(-width // 62)
    return (n // tx, tx) if n else (0, 0)


def visits_buffer(width: int, height: int, lambda_: float, device="cuda", **tiling) -> torch.Tensor:
    """Zeroed int32 per-tile visit counters (tiles_y, tiles_x) for a call with
    the given lambda_ and schedule / tile_rows / tile_waves / max_sweeps; pass
    as round_per_blocks."""
    return torch.zeros(tile_grid(width, height, lambda_, **tiling), dtype=torch.int32, device=device)


def _batch_arrays(imgs, masks, outs):
    n = len(imgs)
    A = DtkImage * max(n, 1)
    ai, am, ad = A(), A(), A()
    for i in range(n):
        ai[i] = _as_image(imgs[i], "image")
        am[i] = _as_image(masks[i], "mask")
        ad[i] = _as_image(outs[i], "dist")
    return ai, am, ad


def ggdt_batched(imgs: Sequence[torch.Tensor], masks: Sequence[torch.Tensor], lambda_: float, v: float = 1e10, *,
                 connectivity: int = 8, tile_rows: int = 0, max_sweeps: int = 0, waves_per_cu: int = 0,
                 outs: Optional[List[torch.Tensor]] = None, stats: bool = False, cost_plane: bool = False,
                 tile_waves: int = 0, queue_order: Optional[str] = None):
    """GGDT of many images of any sizes in one persistent launch."""
    if len(imgs) != len(masks):
        raise ValueError("imgs and masks differ in length")
    if outs is None:
        outs = [torch.empty(m.shape[:2], dtype=torch.float32, device=m.device) for m in masks]
    p = _params(lambda_, v, connectivity, "task", 0, tile_rows, max_sweeps, waves_per_cu,
                cost_plane=cost_plane, tile_waves=tile_waves, queue_order=queue_order)
    ai, am, ad = _batch_arrays(imgs, masks, outs)
    st = DtkStats()
    dev = imgs[0].device if imgs else torch.device("cuda")
    with torch.cuda.device(dev):
        check(_lib.load().dtk_ggdt_batched(len(imgs), ai, am, ad, ctypes.byref(p), _stream_handle(dev),
                                           ctypes.byref(st) if stats else None))
    return (outs, st.as_dict()) if stats else outs


def autotune(imgs, masks, lambda_, v=1e10, *, connectivity=8, outs=None, reps=2, cost_plane=False):
    """Pick tile_rows / max_sweeps / waves_per_cu for a batched plan by timing
    candidate plans on this batch (dtk_batch_autotune).  Returns
    (dict of the chosen parameters, report dict).  Memoised per process."""
    if outs is None:
        outs = [torch.empty(m.shape[:2], dtype=torch.float32, device=m.device) for m in masks]
    p = _params(lambda_, v, connectivity, "task", 0, 0, 0, 0, cost_plane=cost_plane)
    ai, am, ad = _batch_arrays(list(imgs), list(masks), list(outs))
    best, rep = DtkParams(), _lib.DtkTuneReport()
    dev = imgs[0].device
    with torch.cuda.device(dev):
        check(_lib.load().dtk_batch_autotune(len(imgs), ai, am, ad, ctypes.byref(p), _stream_handle(dev), int(reps),
                                             ctypes.byref(best), ctypes.byref(rep)))
    order = {v: k for k, v in _ORDER.items() if k and k != "default"}.get(best.queue_order, "key")
    return ({"tile_rows": best.tile_rows, "tile_waves": best.tile_waves, "max_sweeps": best.max_sweeps,
             "waves_per_cu": best.waves_per_cu, "queue_order": order}, rep.as_dict())


class BatchPlan:
    """A batched GGDT captured as a HIP graph over fixed buffers.

    The tensors given at creation are bound into the graph: refill them in
    place and call ``launch()`` to recompute every ``outs[i]``.  With
    ``tune=True`` the tile height, sweep cap and workers per CU come from
    ``autotune`` (timed on these buffers before capture).
    """

    def __init__(self, imgs, masks, lambda_, v=1e10, *, connectivity=8, tile_rows=0, max_sweeps=0,
                 waves_per_cu=0, outs=None, cost_plane=False, tile_waves=0, tune=False, queue_order=None):
        if outs is None:
            outs = [torch.empty(m.shape[:2], dtype=torch.float32, device=m.device) for m in masks]
        self.imgs, self.masks, self.outs = list(imgs), list(masks), list(outs)
        self.device = self.imgs[0].device
        self.tuning = None
        if tune:
            chosen, report = autotune(self.imgs, self.masks, lambda_, v, connectivity=connectivity, outs=self.outs,
                                      cost_plane=cost_plane)
            tile_rows, max_sweeps, waves_per_cu = chosen["tile_rows"], chosen["max_sweeps"], chosen["waves_per_cu"]
            tile_waves = chosen["tile_waves"]
            queue_order = chosen["queue_order"]
            self.tuning = {**chosen, **report}
        self.params = {"tile_rows": tile_rows, "tile_waves": tile_waves, "max_sweeps": max_sweeps,
                       "waves_per_cu": waves_per_cu, "queue_order": queue_order or "default"}
        p = _params(lambda_, v, connectivity, "task", 0, tile_rows, max_sweeps, waves_per_cu,
                    cost_plane=cost_plane, tile_waves=tile_waves, queue_order=queue_order)
        ai, am, ad = _batch_arrays(self.imgs, self.masks, self.outs)
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            torch.cuda.synchronize(self.device)
            check(_lib.load().dtk_batch_plan_create(len(self.imgs), ai, am, ad, ctypes.byref(p), None,
                                                    ctypes.byref(self._plan)))

    def launch(self):
        with torch.cuda.device(self.device):
            check(_lib.load().dtk_batch_plan_launch(self._plan, _stream_handle(self.device)))
        return self.outs

    def close(self):
        if self._plan:
            _lib.load().dtk_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def predecessors(img, mask, dist, lambda_, v=1e10, *, connectivity=8, out=None) -> torch.Tensor:
    """Argmin predecessor codes (uint8: 0..7 in dt::c8 order, 8 = root)."""
    if out is None:
        out = torch.empty(dist.shape, dtype=torch.uint8, device=dist.device)
    p = _params(lambda_, v, connectivity, "task", 0, 0, 0, 0)
    di, dm, dd, dp = (_as_image(img, "image"), _as_image(mask, "mask"), _as_image(dist, "dist"),
                      _as_image(out, "pred"))
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_ggdt_pred(ctypes.byref(di), ctypes.byref(dm), ctypes.byref(dd), ctypes.byref(dp),
                                        ctypes.byref(p), _stream_handle(img.device)))
    return out


def predecessors_batched(imgs, masks, dists, lambda_, v=1e10, *, connectivity=8, outs=None):
    """Predecessor codes of a batch of converged maps in one launch (images of
    one dtype, masks of one dtype).  Returns the list of uint8 code maps."""
    n = len(imgs)
    if not (len(masks) == len(dists) == n) or (outs is not None and len(outs) != n):
        raise ValueError("predecessors_batched: sequences of different lengths")
    if outs is None:
        outs = [torch.empty(d.shape, dtype=torch.uint8, device=d.device) for d in dists]
    p = _params(lambda_, v, connectivity, "task", 0, 0, 0, 0)
    arr = lambda ts, name: (DtkImage * max(1, n))(*[_as_image(t, name) for t in ts])  # noqa: E731
    if n:
        with torch.cuda.device(imgs[0].device):
            check(_lib.load().dtk_ggdt_pred_batched(n, arr(imgs, "image"), arr(masks, "mask"), arr(dists, "dist"),
                                                    arr(outs, "pred"), ctypes.byref(p),
                                                    _stream_handle(imgs[0].device)))
    return outs


def backward(img, mask, dist, pred, grad_dist, lambda_, v=1e10, *, connectivity=8):
    """Adjoint of sum(grad_dist * D): returns (dL/dimg, dL/dmask)."""
    gi = torch.empty_like(img, dtype=torch.float32)
    gm = torch.empty(mask.shape, dtype=torch.float32, device=mask.device)
    p = _params(lambda_, v, connectivity, "task", 0, 0, 0, 0)
    gd = grad_dist.contiguous()
    ims = [_as_image(t, n) for t, n in ((img, "image"), (mask, "mask"), (dist, "dist"), (pred, "pred"),
                                        (gd, "grad_dist"), (gi, "grad_img"), (gm, "grad_mask"))]
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_ggdt_backward(*[ctypes.byref(x) for x in ims], ctypes.byref(p),
                                            _stream_handle(img.device)))
    return gi, gm


def backward_batched(imgs, masks, dists, preds, grad_dists, lambda_, v=1e10, *, connectivity=8):
    """Adjoint of sum(grad_dists[i] * D_i) for a batch, one launch of each
    adjoint kernel for all images (their predecessor chains climb together).
    Sequences of equal length; returns (list of dL/dimg, list of dL/dmask)."""
    n = len(imgs)
    if not (len(masks) == len(dists) == len(preds) == len(grad_dists) == n):
        raise ValueError("backward_batched: sequences of different lengths")
    gis = [torch.empty_like(t, dtype=torch.float32) for t in imgs]
    gms = [torch.empty(m.shape, dtype=torch.float32, device=m.device) for m in masks]
    gds = [g.contiguous() for g in grad_dists]
    p = _params(lambda_, v, connectivity, "task", 0, 0, 0, 0)
    arr = lambda ts, name: (DtkImage * max(1, n))(*[_as_image(t, name) for t in ts])  # noqa: E731
    A = [arr(imgs, "image"), arr(masks, "mask"), arr(dists, "dist"), arr(preds, "pred"), arr(gds, "grad_dist"),
         arr(gis, "grad_img"), arr(gms, "grad_mask")]
    if n:
        with torch.cuda.device(imgs[0].device):
            check(_lib.load().dtk_ggdt_backward_batched(n, *A, ctypes.byref(p), _stream_handle(imgs[0].device)))
    return gis, gms


def tile_size(lambda_: float, *, schedule: str = "task", tile_rows: int = 0, tile_waves: int = 0,
              max_sweeps: int = 0):
    """(tile_w, tile_h): owned pixels of one tile of a call with these
    parameters (dtk_tile_size_for); pair with ``tile_grid``."""
    p = _params(lambda_, 1e10, 8, schedule, 0, tile_rows, max_sweeps, 0, tile_waves=tile_waves)
    tw, th = ctypes.c_int(), ctypes.c_int()
    if not _lib.load().dtk_tile_size_for(ctypes.byref(p), ctypes.byref(tw), ctypes.byref(th)):
        raise ValueError("invalid tile parameters")
    return tw.value, th.value


def tile_image(values: torch.Tensor, width: int, height: int, tile_w: int, tile_h: int,
               out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Per-tile values (an int32 or uint8 device tensor of shape (tiles_y,
    tiles_x)) spread over their tiles' pixels: a float32 (height, width)
    image, the reference's block_attribute_to_image
    (bin/rounds_per_blocks.cu:13-23) for any tile size (dtk_tile_image)."""
    if values.dim() != 2 or not values.is_cuda or not values.is_contiguous():
        raise ValueError("values must be a contiguous (tiles_y, tiles_x) device tensor")
    if values.dtype not in (torch.int32, torch.uint8):
        raise ValueError("values must be int32 or uint8")
    if out is None:
        out = torch.empty((height, width), dtype=torch.float32, device=values.device)
    nb = 4 if values.dtype == torch.int32 else 1
    with torch.cuda.device(values.device):
        check(_lib.load().dtk_tile_image(ctypes.c_void_p(values.data_ptr()), nb, values.shape[1], values.shape[0],
                                         int(tile_w), int(tile_h), ctypes.byref(_as_image(out, "out")),
                                         _stream_handle(values.device)))
    return out


def heatmap(values: torch.Tensor, width: int, height: int, tile_w: int, tile_h: int) -> torch.Tensor:
    """Inferno heat map (uint8 (height, width, 3)) of per-tile values, e.g.
    the visit counts of ``visits_buffer`` or the priorities of
    ``block_priorities``: tile_image, then normalisation and the colour map
    (the reference's bin/rounds_per_blocks.cu:51-58 and
    bin/priority_block.cu:44-60)."""
    return colorize(tile_image(values, width, height, tile_w, tile_h))


def synchronize(device=None) -> None:
    """Wait for torch's current stream on `device` and raise any error that
    an earlier asynchronous call left on the device (dtk_sync_errors): a
    soft mask value outside [0, 1] raises ValueError, a device fault
    DtkError.  Without it such an error is raised by the next call."""
    dev = torch.device("cuda") if device is None else torch.device(device)
    with torch.cuda.device(dev):
        check(_lib.load().dtk_sync_errors(_stream_handle(dev)))


class KernelTimer:
    """Sums the device time of the task-schedule kernels of every call made
    inside the ``with`` block (HIP events recorded on each call's stream)."""

    def __enter__(self):
        check(_lib.load().dtk_profile_begin())
        return self

    def __exit__(self, *exc):
        a, b, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        check(_lib.load().dtk_profile_end(ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        self.setup_ms, self.solve_ms, self.calls = a.value, b.value, n.value
        return False


def selftest() -> dict:
    out = (ctypes.c_int * 3)()
    check(_lib.load().dtk_selftest(out))
    return {"dpp_ok": bool(out[0]), "sqrt_mismatches": int(out[1]), "cu": int(out[2])}


# ---- display of a distance map (the reference CLI's last steps) -------------
def minmax(img: torch.Tensor):
    """(min, max) of a float32 device image as dt::minmax computes them
    (dt/include/dt/normalize.hpp:28-43): NaNs skipped, max starting at the
    smallest positive float.  Synchronous."""
    out = (ctypes.c_float * 2)()
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_minmax_f32(ctypes.byref(_as_image(img, "img")), out, _stream_handle(img.device)))
    return float(out[0]), float(out[1])


def normalize_u8(img: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """float32 (H, W) -> uint8 (H, W): (uint8)(double(x - min) / double(max - min) * 255),
    truncated (dt/include/dt/normalize.hpp:45-62; 0/0 gives 0)."""
    if out is None:
        out = torch.empty(img.shape[:2], dtype=torch.uint8, device=img.device)
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_normalize_u8(ctypes.byref(_as_image(img, "img")), ctypes.byref(_as_image(out, "out")),
                                           _stream_handle(img.device)))
    return out


def inferno(img: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """uint8 (H, W) -> uint8 (H, W, 3) through the inferno colour map (dt/src/inferno.cpp:263-274)."""
    if out is None:
        out = torch.empty((img.shape[0], img.shape[1], 3), dtype=torch.uint8, device=img.device)
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_inferno(ctypes.byref(_as_image(img, "img")), ctypes.byref(_as_image(out, "out")),
                                      _stream_handle(img.device)))
    return out


def block_priorities(mask: torch.Tensor):
    """Tile priorities of the reference's task queue, dt::compute_block_priorities
    (dt/include/dt/block_priority.hpp:7-8, dt/src/block_priority.cu:74-155), for
    a uint8 or float32 device mask (nonzero = seed) cut into 32 x 32 tiles.

    Returns (priorities, offsets, level0): priorities uint8 (ceil(H/32),
    ceil(W/32)) = cdf[d] * 64 // (N + 1), with d the city-block distance in
    tiles to the nearest seed tile and cdf[d] the number of tiles at distance
    <= d; offsets int32 (64,), the number of tiles of priority < k; level0
    the number of seed tiles.  Synchronous."""
    H, W = mask.shape[0], mask.shape[1]
    gh, gw = (H + 31) // 32, (W + 31) // 32
    prio = torch.empty((gh, gw), dtype=torch.uint8, device=mask.device)
    offsets = torch.zeros(64, dtype=torch.int32, device=mask.device)
    level0 = ctypes.c_uint32(0)
    with torch.cuda.device(mask.device):
        check(_lib.load().dtk_block_priorities(ctypes.byref(_as_image(mask, "mask")), ctypes.c_void_p(prio.data_ptr()),
                                               ctypes.c_void_p(offsets.data_ptr()), ctypes.byref(level0),
                                               _stream_handle(mask.device)))
    return prio, offsets, int(level0.value)


def colorize(img: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """float32 (H, W) -> uint8 (H, W, 3): normalize_u8 then inferno in one pass."""
    if out is None:
        out = torch.empty((img.shape[0], img.shape[1], 3), dtype=torch.uint8, device=img.device)
    with torch.cuda.device(img.device):
        check(_lib.load().dtk_colorize(ctypes.byref(_as_image(img, "img")), ctypes.byref(_as_image(out, "out")),
                                       _stream_handle(img.device)))
    return out
