"""HBM traffic of a kernel from rocprofv3 PMC counters.

Three sub-commands:

  copy       one device-to-device copy of a known size (torch ``copy_`` of a
             float32 tensor), the calibration workload: its FETCH_SIZE and
             WRITE_SIZE are compared with the bytes it must move;
  report     writes a text summary of a kernel-trace run, the bench line of
             the same run and a traffic summary (the files under profiles/);
  summarize  reads the rocprofv3 ``counter_collection.csv`` files of a
             FETCH_SIZE run and a WRITE_SIZE run of the workload and of the
             copy, scales each counter by the copy's known/measured ratio and
             writes a JSON summary with the HBM bytes per dispatch of one
             kernel.

FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch.  They need
separate runs (FETCH_SIZE takes three TCC counters, WRITE_SIZE two).

Example (each rocprofv3 line is its own run):

  rocprofv3 --pmc FETCH_SIZE -d out/f -o run --output-format csv -- python3 bench.py --steps 2 --warmup 1 --no-cpu --no-batched
  rocprofv3 --pmc WRITE_SIZE -d out/w -o run --output-format csv -- python3 bench.py --steps 2 --warmup 1 --no-cpu --no-batched
  rocprofv3 --pmc FETCH_SIZE -d out/cf -o run --output-format csv -- python3 -m distance_transform_benchmarks_amd.traffic copy
  rocprofv3 --pmc WRITE_SIZE -d out/cw -o run --output-format csv -- python3 -m distance_transform_benchmarks_amd.traffic copy
  python3 -m distance_transform_benchmarks_amd.traffic summarize --kernel k_task --fetch out/f --write out/w \\
      --copy-fetch out/cf --copy-write out/cw --algorithmic-bytes 184549376 --out traffic.json
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

COPY_BYTES = 1 << 30  # bytes read (and written) by the calibration copy


def run_copy(nbytes: int = COPY_BYTES, reps: int = 3) -> None:
    import torch

    dev = torch.device("cuda:0")
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy: {reps} x {nbytes} bytes", flush=True)


def _csv_rows(path: str):
    files = [path] if os.path.isfile(path) else sorted(
        glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {path}")
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def per_dispatch(path: str, counter: str, kernel_substr: str):
    """Counter value (KiB) of every dispatch whose kernel name contains kernel_substr."""
    vals = {}
    for r in _csv_rows(path):
        name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
        if kernel_substr not in name:
            continue
        if (r.get("Counter_Name") or "").strip() != counter:
            continue
        key = r.get("Dispatch_Id") or r.get("Correlation_Id") or str(len(vals))
        vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    return list(vals.values())


def _mean(v):
    return sum(v) / len(v) if v else float("nan")


def summarize(a) -> dict:
    kib = 1024.0
    f = per_dispatch(a.fetch, "FETCH_SIZE", a.kernel)
    w = per_dispatch(a.write, "WRITE_SIZE", a.kernel)
    cf = per_dispatch(a.copy_fetch, "FETCH_SIZE", a.copy_kernel)
    cw = per_dispatch(a.copy_write, "WRITE_SIZE", a.copy_kernel)
    if not (f and w and cf and cw):
        raise SystemExit(f"missing dispatches: {len(f)} {len(w)} {len(cf)} {len(cw)}")
    # The calibration copy moves exactly COPY_BYTES each way; take the largest
    # dispatches (the copy itself, not small fills).
    cf_b, cw_b = max(cf) * kib, max(cw) * kib
    kf, kw = a.copy_bytes / cf_b, a.copy_bytes / cw_b
    fetch_b, write_b = _mean(f) * kib * kf, _mean(w) * kib * kw
    out = {
        "kernel": a.kernel, "dispatches_fetch": len(f), "dispatches_write": len(w),
        "fetch_bytes_per_launch": round(fetch_b), "write_bytes_per_launch": round(write_b),
        "bytes_per_launch": round(fetch_b + write_b),
        "raw_fetch_kib_mean": round(_mean(f), 1), "raw_write_kib_mean": round(_mean(w), 1),
        "copy_calibration": {"bytes": a.copy_bytes, "fetch_kib": max(cf), "write_kib": max(cw),
                             "fetch_scale": round(kf, 4), "write_scale": round(kw, 4)},
    }
    if a.algorithmic_bytes:
        out["algorithmic_bytes_per_launch"] = a.algorithmic_bytes
        out["traffic_over_algorithmic"] = round((fetch_b + write_b) / a.algorithmic_bytes, 3)
    return out


def report(a) -> str:
    """Text summary of a kernel-trace run (rocprofv3 --kernel-trace --stats
    kernel_stats.csv), the bench line of the same run and a traffic summary."""
    files = sorted(glob.glob(os.path.join(a.trace, "**", "*kernel_stats.csv"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no kernel_stats.csv under {a.trace}")
    rows = list(csv.DictReader(open(files[0], newline="")))
    bench = json.load(open(a.bench)) if a.bench else None
    traffic = json.load(open(a.traffic)) if a.traffic else None
    out = [f"# {a.title}", "", "Kernel statistics (rocprofv3 --kernel-trace --stats; every dispatch of the run):", "",
           f"{'kernel':<70} {'calls':>5} {'avg us':>10} {'min us':>10} {'max us':>10} {'%':>6}"]
    for r in rows:
        out.append(f"{r['Name'][:70]:<70} {r['Calls']:>5} {float(r['AverageNs']) / 1e3:>10.1f} "
                   f"{float(r['MinNs']) / 1e3:>10.1f} {float(r['MaxNs']) / 1e3:>10.1f} {float(r['Percentage']):>6.2f}")
    kt = [r for r in rows if a.kernel in r["Name"]]
    if kt:
        tot = sum(float(r["TotalDurationNs"]) for r in kt)
        n = sum(int(r["Calls"]) for r in kt)
        out += ["", f"{a.kernel}: {n} dispatches, average {tot / n / 1e6:.4f} ms"]
    if a.split_us > 0:
        # Two workloads in one run that launch the same kernel (a long solve
        # and the short ones around it) are told apart by duration.
        traces = sorted(glob.glob(os.path.join(a.trace, "**", "*kernel_trace.csv"), recursive=True))
        if not traces:
            raise FileNotFoundError(f"no kernel_trace.csv under {a.trace}")
        durs = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                for r in csv.DictReader(open(traces[0], newline="")) if a.kernel in r["Kernel_Name"]]
        for label, sel in (("at least", [d for d in durs if d >= a.split_us]),
                           ("under", [d for d in durs if d < a.split_us])):
            if sel:
                out.append(f"  {len(sel)} dispatches {label} {a.split_us:g} us: average {sum(sel) / len(sel) / 1e3:.4f} ms, "
                           f"min {min(sel) / 1e3:.4f}, max {max(sel) / 1e3:.4f}")
    if bench:
        rl = bench["roofline"]
        out += [f"bench line of the same run: {a.kernel} average from HIP events on its stream {rl['kernel_avg_ms']} ms "
                f"over {rl['kernel_launches']} timed launches -> algorithmic {rl['achieved']} GB/s = "
                f"{rl['frac']} of {rl['peak']} GB/s ({rl['algorithmic_bytes_per_launch']} B per launch)",
                f"value {bench['value']} {bench['unit']}, {bench['ms_per_step']} ms per step"]
    if traffic:
        c = traffic["copy_calibration"]
        out += ["", "HBM traffic per dispatch (FETCH_SIZE and WRITE_SIZE in separate runs, each scaled by its reading "
                "for a device copy of known size):",
                f"  copy of {c['bytes']} B: FETCH_SIZE {c['fetch_kib']} KiB (scale {c['fetch_scale']}), "
                f"WRITE_SIZE {c['write_kib']} KiB (scale {c['write_scale']})",
                f"  {traffic['kernel']}: raw FETCH {traffic['raw_fetch_kib_mean']} KiB, WRITE "
                f"{traffic['raw_write_kib_mean']} KiB per dispatch ({traffic['dispatches_fetch']} dispatches)",
                f"  scaled: fetch {traffic['fetch_bytes_per_launch'] / 1e6:.1f} MB + write "
                f"{traffic['write_bytes_per_launch'] / 1e6:.1f} MB = {traffic['bytes_per_launch'] / 1e6:.1f} MB per launch"]
        if "traffic_over_algorithmic" in traffic:
            out.append(f"  = {traffic['traffic_over_algorithmic']} x the algorithmic bytes per launch")
    if a.commands:
        out += ["", "Commands (repository root, TMPDIR=/tmp):"] + [f"  {c}" for c in a.commands]
    return "\n".join(out) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--bytes", type=int, default=COPY_BYTES)
    s = sub.add_parser("summarize")
    s.add_argument("--kernel", default="k_task")
    s.add_argument("--copy-kernel", default="")
    s.add_argument("--fetch", required=True)
    s.add_argument("--write", required=True)
    s.add_argument("--copy-fetch", required=True)
    s.add_argument("--copy-write", required=True)
    s.add_argument("--copy-bytes", type=int, default=COPY_BYTES)
    s.add_argument("--algorithmic-bytes", type=int, default=0)
    s.add_argument("--out", default="")
    r = sub.add_parser("report")
    r.add_argument("--trace", required=True)
    r.add_argument("--kernel", default="k_task")
    r.add_argument("--bench", default="")
    r.add_argument("--traffic", default="")
    r.add_argument("--title", default="kernel trace")
    r.add_argument("--split-us", type=float, default=0.0,
                   help="also report the kernel's dispatches at least / under this many us apart (kernel_trace.csv)")
    r.add_argument("--commands", nargs="*", default=[])
    r.add_argument("--out", default="")
    a = ap.parse_args(argv)
    if a.cmd == "report":
        txt = report(a)
        if a.out:
            with open(a.out, "w") as fh:
                fh.write(txt)
        print(txt, end="")
        return 0
    if a.cmd == "copy":
        run_copy(a.bytes)
        return 0
    res = summarize(a)
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(txt + "\n")
    print(txt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
