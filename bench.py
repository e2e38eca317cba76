#!/usr/bin/env python3
"""Benchmark of the MI355X GGDT on its headline workload.

Metric: converged Mpix/s on one 4096x4096 RGB8 image (three seeded blob
fields), soft float32 seed mask of 64 discs (nu = 1e3), C8, lambda swept over
{0.0, 0.5, 1.0}; the schedule is the persistent tile queue ("task").  One
step = the three converged transforms of the sweep; inputs and outputs stay
resident in HBM; the timed region is bracketed by device synchronisation.

The JSON line also carries:
  roofline      the persistent solve kernel (k_task): algorithmic bytes per
                launch (11 B/px: RGB8 3 + D read 4 + D write 4) over its
                average duration, from HIP events recorded on its stream
                during the timed steps; `traffic` comes from the rocprofv3
                FETCH_SIZE / WRITE_SIZE summary committed under profiles/
                (null when absent); operator_frac is the same figure for
                the whole transform (the setup kernels' time added; its
                compulsory bytes are also 11 B/px: image 3 + mask 4 +
                D write 4);
  cpu_baseline  the CPU oracle (a restatement of the reference's raster GGDT)
                on the same 4096x4096 image at lambda=0.5, OpenMP over the
                box's cores (at most 16), median of 5 runs after a warm-up,
                also checked bit-exact against the GPU result; with the
                core counts, and a single-thread run on blob512;
  orders        per lambda, ms and visits of each queue order (FIFO, KEY,
                SEED) on the same image;
  maze          the 8192x8192 spiral maze (C8, lambda=1): ms per converged
                call, Mpix/s, visits per tile per queue order, D == I;
  blob512       the reference-CPU-sized workload (512x512 u8, C4,
                lambda=1): GPU ms (median of 10) and the oracle;
  batched       256 uneven uint8 images (C8, lambda=0.8): one graph-captured
                persistent launch vs 256 per-image calls, images/s, and
                visits per queue order;
  display       the reference CLI's step after the transform on the
                lambda=0.5 map: min / max, normalisation to uint8 and the
                inferno colour map (dtk_colorize, two kernels), timed with
                HIP events; algorithmic 11 B/px (min/max read 4, colour pass
                read 4 + write 3) against the HBM peak.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged Mpix/s @4096² RGB8 C8 + % HBM roofline; batched vs per-image imgs/s"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PX = 11  # RGB8 3 + D read 4 + D write 4 (k_task); mask 4 B/px is read by k_init


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def load_traffic():
    """HBM bytes per k_task launch measured by rocprofv3 (profiles/traffic_k_task_rgb4096.json)."""
    p = os.path.join(ROOT, "profiles", "traffic_k_task_rgb4096.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


# CPU threads of one GPU's share of the box (worker pools are sized to it).
CPU_SHARE = 16


def cpu_counts():
    """(hardware threads, affinity cores, threads used)."""
    hw = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = hw
    return hw, aff, max(1, min(CPU_SHARE, aff))


def median_time(fn, runs=5):
    """Median wall time (monotonic clock) of `runs` calls after one warm-up."""
    fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    ts.sort()
    return ts[len(ts) // 2], ts


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--max-sweeps", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--tile-waves", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable (adjoint) leg")
    ap.add_argument("--no-maze", action="store_true", help="skip the maze leg")
    ap.add_argument("--no-orders", action="store_true", help="skip the per-lambda queue-order comparison")
    ap.add_argument("--no-blob512", action="store_true", help="skip the blob512 leg (and its single-thread CPU run)")
    ap.add_argument("--batch-count", type=int, default=256)
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU (no multi-GPU path); use --gpus 1"}))
        return 2

    import torch

    import distance_transform_benchmarks_amd as dt
    from distance_transform_benchmarks_amd import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lambdas = (0.0, 0.5, 1.0)
    kw = dict(connectivity=8, tile_rows=a.tile_rows, max_sweeps=a.max_sweeps, waves_per_cu=a.waves_per_cu,
              tile_waves=a.tile_waves)

    t0 = time.time()
    img_np, mask_np = synth.rgb4096(a.size)
    log(f"inputs {a.size}x{a.size} RGB8 generated in {time.time() - t0:.1f}s")
    img = torch.from_numpy(img_np).to(dev)
    mask = torch.from_numpy(mask_np).to(dev)
    outs = {lam: torch.empty((a.size, a.size), dtype=torch.float32, device=dev) for lam in lambdas}
    npx = a.size * a.size

    def step():
        for lam in lambdas:
            dt.ggdt(img, mask, lam, 1e3, out=outs[lam], **kw)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()

    # Per-lambda statistics (outside the timed region; these calls synchronise),
    # and the same call under each queue order.
    per_lambda = {}
    for lam in lambdas:
        _, st = dt.ggdt(img, mask, lam, 1e3, out=outs[lam], stats=True, **kw)
        ms = event_ms(torch, lambda: dt.ggdt(img, mask, lam, 1e3, out=outs[lam], **kw), 3)
        per_lambda[str(lam)] = {"ms": round(ms, 4), "mpix_s": round(npx / ms / 1e3, 1), "visits": st["visits"],
                                "sweeps": st["sweeps"], "tiles": st["tiles"], "tile_rows": st["tile_rows"],
                                "workers": st["workers"]}
        log(f"lambda={lam}: {ms:.3f} ms/call, {per_lambda[str(lam)]['mpix_s']} Mpix/s, stats {st}")
        if a.no_orders:
            continue
        orders = {}
        for order in ("fifo", "key", "seed"):
            okw = dict(kw, queue_order=order)
            _, so = dt.ggdt(img, mask, lam, 1e3, out=outs[lam], stats=True, **okw)
            oms = event_ms(torch, lambda: dt.ggdt(img, mask, lam, 1e3, out=outs[lam], **okw), 3)
            orders[order] = {"ms": round(oms, 4), "visits": so["visits"],
                             "sweeps_per_visit": round(so["sweeps"] / max(1, so["visits"]), 2),
                             "visits_per_tile": round(so["visits"] / max(1, so["tiles"]), 3)}
        per_lambda[str(lam)]["orders"] = orders
        log(f"lambda={lam} orders: {orders}")

    # Timed region.
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with dt.api.KernelTimer() as kt:
        e0.record()
        for _ in range(a.steps):
            step()
        e1.record()
        torch.cuda.synchronize()
    elapsed_ms = e0.elapsed_time(e1)
    ms_per_step = elapsed_ms / a.steps
    value = a.steps * len(lambdas) * npx / (elapsed_ms * 1e3)
    solve_avg_ms = kt.solve_ms / max(1, kt.calls)
    setup_avg_ms = kt.setup_ms / max(1, kt.calls)
    achieved = BYTES_PER_PX * npx / (solve_avg_ms * 1e-3) / 1e9
    traffic = load_traffic()
    roofline = {
        "bound": "hbm", "kernel": "k_task (persistent tile solve)", "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic.get("bytes_per_launch") if traffic else None,
        "algorithmic_bytes_per_launch": BYTES_PER_PX * npx, "kernel_avg_ms": round(solve_avg_ms, 4),
        "setup_avg_ms": round(setup_avg_ms, 4), "kernel_launches": kt.calls,
        # The whole transform (reset, init, seed and solve): its compulsory
        # bytes (image 3 + mask 4 + D write 4 = 11 B/px) over setup + solve.
        "operator_achieved": round(BYTES_PER_PX * npx / ((solve_avg_ms + setup_avg_ms) * 1e-3) / 1e9, 2),
        "operator_frac": round(BYTES_PER_PX * npx / ((solve_avg_ms + setup_avg_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
    }
    log(f"timed: {ms_per_step:.3f} ms/step, {value:.1f} Mpix/s; k_task avg {solve_avg_ms:.3f} ms "
        f"-> {achieved:.1f} GB/s algorithmic")

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded SplitMix64 blob fields and discs)",
        "config": {"workload": f"rgb4096: {a.size}x{a.size} RGB8, soft f32 mask (64 discs, nu=1e3), C8, "
                               f"lambda sweep {{0, 0.5, 1}}, converged, task schedule",
                   "pixels_per_step": len(lambdas) * npx, "tile_rows": per_lambda["0.0"]["tile_rows"]},
        "per_lambda": per_lambda, "roofline": roofline,
    }

    result["display"] = bench_display(dt, torch, outs[0.5], npx)

    if not a.no_batched:
        result["batched"] = bench_batched(dt, synth, torch, dev, a.batch_count)

    if not a.no_adjoint:
        result["adjoint"] = bench_adjoint(dt, synth, torch, dev)

    if not a.no_maze:
        result["maze"] = bench_maze(dt, synth, torch, dev)

    if not a.no_blob512:
        result["blob512"] = bench_blob512(dt, synth, torch, dev, a.no_cpu)

    if not a.no_cpu:
        from oracle import oracle as O

        hw, aff, threads = cpu_counts()
        lam = 0.5
        res = {}

        def run():
            res["D"], res["iters"] = O.raster(img_np, mask_np, lam, 1e3, conn8=True, threads=threads)

        cpu_s, runs = median_time(run, 5)
        got = outs[lam].cpu().numpy()
        exact = bool(np.array_equal(got.view(np.uint32), res["D"].view(np.uint32)))
        result["cpu_baseline"] = {
            "value": round(npx / cpu_s / 1e6, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle raster GGDT, full {a.size}x{a.size} RGB8 image at lambda=0.5 ({res['iters']} raster "
                      f"iterations; median of 5 runs after a warm-up: {cpu_s:.2f}s; OpenMP {threads} threads)",
            "hardware_concurrency": hw, "affinity_cores": aff, "threads": threads,
            "threads_note": f"OpenMP over the affinity cores, at most {CPU_SHARE} (one GPU's share of the box)",
            "runs_s": [round(t, 3) for t in runs], "timer": "monotonic wall clock (time.perf_counter)",
            "gpu_bit_exact": exact,
            "single_thread": result.get("blob512", {}).get("cpu_single_thread"),
        }
        log(f"cpu baseline: {cpu_s:.2f}s median ({res['iters']} iterations, {threads} of {aff} affinity cores, "
            f"{hw} hardware threads); GPU bit-exact: {exact}")

    print(json.dumps(result), flush=True)
    return 0


def event_ms(torch, fn, reps):
    """Mean ms per call of `fn` between two events on the current stream."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def bench_maze(dt, synth, torch, dev):
    """The adversarial spiral maze: 8192^2 u8, walls 255, one seed in the
    corridor, C8, lambda=1; the known answer is D = I.  6 B/px algorithmic
    (image 1 + mask 1 + D 4)."""
    t = time.time()
    img_np, mask_np = synth.maze8192()
    img, mask = torch.from_numpy(img_np).to(dev), torch.from_numpy(mask_np).to(dev)
    out = torch.empty(img.shape, dtype=torch.float32, device=dev)
    npx = img.numel()
    log(f"maze inputs generated in {time.time() - t:.1f}s")
    dt.ggdt(img, mask, 1.0, 1e10, out=out)
    ms = event_ms(torch, lambda: dt.ggdt(img, mask, 1.0, 1e10, out=out), 2)
    exact = bool(torch.equal(out, img.float()))
    gbs = 6 * npx / (ms * 1e-3) / 1e9
    r = {"workload": "maze8192: 8192x8192 u8 square spiral maze (walls 3 px = 255, corridor 9 px = 0), one seed "
                     "at the centre, C8, lambda=1, converged, task schedule (default order)",
         "ms": round(ms, 2), "mpix_s": round(npx / ms / 1e3, 1), "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 6), "algorithmic_bytes": 6 * npx, "D_equals_I": exact,
         "orders": {}}
    for order in ("fifo", "key", "seed"):
        t0 = time.time()
        _, st = dt.ggdt(img, mask, 1.0, 1e10, out=out, stats=True, queue_order=order)
        r["orders"][order] = {"ms": round((time.time() - t0) * 1e3, 1), "visits": st["visits"],
                              "visits_per_tile": round(st["visits"] / max(1, st["tiles"]), 1)}
    log(f"maze: {ms:.1f} ms ({r['mpix_s']} Mpix/s), D == I: {exact}, orders {r['orders']}")
    return r


def bench_blob512(dt, synth, torch, dev, no_cpu):
    """The reference-CPU-sized workload: 512^2 u8 blob field, 8 seed discs,
    C4, lambda=1: GPU time (median of 10 calls) and, unless no_cpu, the
    single-thread oracle (median of 5 runs after a warm-up)."""
    img_np, mask_np = synth.blob512()
    img, mask = torch.from_numpy(img_np).to(dev), torch.from_numpy(mask_np).to(dev)
    out = torch.empty(img.shape, dtype=torch.float32, device=dev)
    npx = img.numel()
    for _ in range(3):
        dt.ggdt(img, mask, 1.0, 1e10, connectivity=4, out=out)
    torch.cuda.synchronize()
    ts = sorted(event_ms(torch, lambda: dt.ggdt(img, mask, 1.0, 1e10, connectivity=4, out=out), 1)
                for _ in range(10))
    ms = ts[len(ts) // 2]
    r = {"workload": "blob512: 512x512 u8 blob field, 8 seed discs (r=5), C4, lambda=1, converged, task schedule",
         "ms": round(ms, 4), "mpix_s": round(npx / ms / 1e3, 1)}
    if not no_cpu:
        from oracle import oracle as O

        res = {}

        def run():
            res["D"], res["iters"] = O.raster(img_np, mask_np, 1.0, 1e10, conn8=False, threads=1)

        cpu_s, runs = median_time(run, 5)
        r["gpu_bit_exact"] = bool(np.array_equal(out.cpu().numpy().view(np.uint32), res["D"].view(np.uint32)))
        r["cpu_single_thread"] = {"value": round(npx / cpu_s / 1e6, 3), "unit": "Mpix/s", "threads": 1,
                                  "median_s": round(cpu_s, 4), "runs_s": [round(t, 4) for t in runs],
                                  "sample": f"CPU oracle, blob512 ({res['iters']} raster iterations), one thread"}
    log(f"blob512: {ms:.4f} ms GPU; {r.get('cpu_single_thread')}")
    return r


def bench_display(dt, torch, D, npx, reps=20):
    """dtk_colorize (min / max + normalise + inferno) of one 4096^2 map."""
    rgb = torch.empty((D.shape[0], D.shape[1], 3), dtype=torch.uint8, device=D.device)
    for _ in range(3):
        dt.colorize(D, out=rgb)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dt.colorize(D, out=rgb)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 11 * npx / (ms * 1e-3) / 1e9
    log(f"display: colorize {ms:.4f} ms per 4096^2 map, {gbs:.0f} GB/s algorithmic")
    return {"workload": "dtk_colorize of the lambda=0.5 map (min/max, normalise to uint8, inferno)", "ms": round(ms, 4),
            "mpix_s": round(npx / ms / 1e3, 1), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": 11 * npx}


def bench_adjoint(dt, synth, torch, dev, count=32, size=1024, reps=3):
    """The differentiable workload: 32 float32 1024^2 blob fields with soft
    masks, C8, lambda=1, loss = sum(W * D).  Forward = batched converged
    solve + predecessor codes; backward = the adjoint kernels, which read
    pred 1 + dL/dD 4 + image 4 B/px and write dL/dimage 4 + dL/dmask 4 B/px."""
    t = time.time()
    imgs, masks, wts = synth.diff32(count, size)
    ti = torch.from_numpy(imgs).to(dev).requires_grad_(True)
    tm = torch.from_numpy(masks).to(dev).requires_grad_(True)
    tw = torch.from_numpy(wts).to(dev)
    log(f"adjoint inputs: {count} x {size}^2, generated in {time.time() - t:.1f}s")
    fwd, bwd = [], []
    for r in range(reps + 1):
        ti.grad = None
        tm.grad = None
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        torch.cuda.synchronize()
        e[0].record()
        D = dt.ggdt_differentiable(ti, tm, 1.0, 1e3)
        loss = (D * tw).sum()
        e[1].record()
        loss.backward()
        e[2].record()
        torch.cuda.synchronize()
        if r:  # the first pass warms up
            fwd.append(e[0].elapsed_time(e[1]))
            bwd.append(e[1].elapsed_time(e[2]))
    f_ms, b_ms = min(fwd), min(bwd)
    npx = count * size * size
    gbs = 17 * npx / (b_ms * 1e-3) / 1e9
    finite = bool(torch.isfinite(ti.grad).all() and torch.isfinite(tm.grad).all())
    log(f"adjoint: forward {f_ms:.2f} ms, backward {b_ms:.2f} ms ({gbs:.0f} GB/s algorithmic), finite {finite}")
    return {"workload": f"diff32: {count} float32 {size}x{size} blob fields + N(0, 1e-3) noise, soft masks, C8, "
                        f"lambda=1, loss = sum(W * D), forward + backward through torch.autograd",
            "pixels": npx, "forward_ms": round(f_ms, 3), "backward_ms": round(b_ms, 3),
            "imgs_s": round(count / (f_ms + b_ms) * 1e3, 1),
            "backward_achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "backward_frac": round(gbs / HBM_PEAK_GBS, 4), "backward_algorithmic_bytes": 17 * npx,
            "grads_finite": finite}


def bench_batched(dt, synth, torch, dev, count):
    """256 uneven uint8 images, C8, lambda=0.8: graph-captured batched launch vs per-image calls."""
    t = time.time()
    data = synth.batch256(count)
    imgs = [torch.from_numpy(i).to(dev) for i, _ in data]
    masks = [torch.from_numpy(m).to(dev) for _, m in data]
    outs_a = [torch.empty(m.shape, dtype=torch.float32, device=dev) for m in masks]
    outs_b = [torch.empty(m.shape, dtype=torch.float32, device=dev) for m in masks]
    px = sum(int(m.numel()) for m in masks)
    log(f"batched inputs: {count} images, {px / 1e6:.1f} Mpx, generated in {time.time() - t:.1f}s")
    lam = 0.8

    def per_image():
        for i in range(count):
            dt.ggdt(imgs[i], masks[i], lam, 1e10, out=outs_a[i])

    t = time.time()
    plan = dt.BatchPlan(imgs, masks, lam, 1e10, outs=outs_b, tune=True)
    log(f"batched plan autotuned in {time.time() - t:.1f}s: {plan.tuning}")
    for _ in range(2):
        per_image()
        plan.launch()
    torch.cuda.synchronize()
    reps = 3
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    for _ in range(reps):
        per_image()
    e[1].record()
    e[2].record()
    for _ in range(reps):
        plan.launch()
    e[3].record()
    torch.cuda.synchronize()
    per_ms = e[0].elapsed_time(e[1]) / reps
    bat_ms = e[2].elapsed_time(e[3]) / reps
    same = all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(outs_a, outs_b))
    plan.close()
    # One launch per queue order (ggdt_batched, not captured): ms and visits.
    orders = {}
    for order in ("fifo", "key", "seed"):
        _, st = dt.ggdt_batched(imgs, masks, lam, 1e10, outs=outs_a, stats=True, queue_order=order)
        oms = event_ms(torch, lambda: dt.ggdt_batched(imgs, masks, lam, 1e10, outs=outs_a, queue_order=order), 2)
        orders[order] = {"ms": round(oms, 3), "visits": st["visits"],
                         "visits_per_tile": round(st["visits"] / max(1, st["tiles"]), 3),
                         "sweeps_per_visit": round(st["sweeps"] / max(1, st["visits"]), 2)}
    log(f"batched orders: {orders}")
    r = {"workload": f"batch256: {count} uint8 images, sides log-uniform [64, 2048], aspect [0.5, 2], C8, "
                     f"lambda=0.8", "pixels": px,
         "per_image_ms": round(per_ms, 3), "per_image_imgs_s": round(count / per_ms * 1e3, 1),
         "batched_graph_ms": round(bat_ms, 3), "batched_imgs_s": round(count / bat_ms * 1e3, 1),
         "speedup": round(per_ms / bat_ms, 3), "outputs_identical": bool(same), "orders": orders,
         "autotuned": {k: plan.tuning[k] for k in ("queue_order", "tile_rows", "tile_waves", "max_sweeps",
                                                   "waves_per_cu", "candidates", "best_ms", "default_ms")}}
    log(f"batched: per-image {per_ms:.2f} ms, graph {bat_ms:.2f} ms, x{per_ms / bat_ms:.2f}, identical {same}")
    return r


if __name__ == "__main__":
    sys.exit(main())
