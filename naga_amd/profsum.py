"""Summaries of rocprofv3 CSV output for the committed profiles.

    python -m naga_amd.profsum stats DIR [--match TEXT] [--title TEXT]
        kernel table from DIR/*_kernel_stats.csv (rocprofv3 --kernel-trace
        --stats --output-format csv)

    python -m naga_amd.profsum traffic --fetch DIR --write DIR --kernel TEXT
        HBM bytes per dispatch of the kernels whose name contains TEXT, from
        two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch)
        over the same program.  Each counter is divided by what it reads,
        per byte, for the runtime's device-to-device copy kernel
        (__amd_rocclr_copyBuffer) of known size in the same directory
        (``python bench.py --copy-calibration BYTES`` under the same pass,
        written with ``-o copy``: only copy_* files calibrate).

Output is Markdown on stdout; the traffic summary carries one line
``traffic_bytes_per_launch = <int>`` that bench.py reads.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def short_name(name: str) -> str:
    name = name.replace("naga_amd::(anonymous namespace)::", "")
    if name.endswith(")"):  # drop the parameter list
        depth = 0
        for i in range(len(name) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(name[i], 0)
            if depth == 0:
                name = name[:i]
                break
    name = name.replace("void ", "")
    return name


def _rows(pattern: str):
    files = sorted(glob.glob(pattern))
    if not files:
        raise SystemExit(f"no file matches {pattern}")
    for path in files:
        with open(path, newline="") as fh:
            yield from csv.DictReader(fh)


def stats(args) -> None:
    rows = [r for r in _rows(os.path.join(args.dir, "*kernel_stats.csv"))
            if args.match in r["Name"]]
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    print(f"# {args.title}\n")
    print("| kernel | calls | avg us | total ms | % |")
    print("|---|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        t = float(r["TotalDurationNs"])
        print(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} "
              f"| {t / 1e6:.3f} | {100 * t / total:.1f} |")


def _per_dispatch(dirname: str, counter: str, match: str):
    """Counter bytes of the dispatches of `match`, and of the copy kernel in
    the calibration run's output (files named copy_*; the program measured
    may run copies of its own)."""
    vals, copies = [], []
    for path in sorted(glob.glob(os.path.join(dirname, "*counter_collection.csv"))):
        calib = os.path.basename(path).startswith("copy")
        with open(path, newline="") as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                v = float(r["Counter_Value"]) * 1024.0  # KiB -> B
                if COPY_KERNEL in r["Kernel_Name"]:
                    if calib:
                        copies.append(v)
                elif match in r["Kernel_Name"]:
                    vals.append(v)
    return vals, copies


def traffic(args) -> None:
    out = {}
    lines = []
    for counter, d in (("FETCH_SIZE", args.fetch), ("WRITE_SIZE", args.write)):
        vals, copies = _per_dispatch(d, counter, args.kernel)
        if not vals or not copies:
            raise SystemExit(f"{counter}: no dispatches of '{args.kernel}' or no copy kernel")
        factor = max(copies) / args.copy_bytes
        raw = sum(vals) / len(vals)
        out[counter] = raw / factor
        lines.append(f"| {counter} | {len(vals)} | {raw / 1e6:.1f} | {factor:.3f} "
                     f"| {out[counter] / 1e6:.1f} |")
    total = out["FETCH_SIZE"] + out["WRITE_SIZE"]
    print(f"# HBM traffic of `{args.kernel}`\n")
    print("| counter | dispatches | raw MB/dispatch | counter bytes per copied byte "
          "| calibrated MB/dispatch |")
    print("|---|---:|---:|---:|---:|")
    print("\n".join(lines))
    print()
    print(f"traffic_bytes_per_launch = {int(round(total))}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--match", default="")
    s.add_argument("--title", default="rocprofv3 kernel statistics")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, default=float(256 << 20))
    args = ap.parse_args(argv)
    {"stats": stats, "traffic": traffic}[args.cmd](args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
