"""Summaries of rocprofv3 CSV output for a timed region of this library's runs.

A bench run spins its domain up before the timed steps, so whole-run kernel statistics mix the
two. The timed region is recovered from the trace: it is the window spanned by the last N
dispatches of the dominant kernel (N = the launches the bench reports), and every kernel's
statistics are taken over that window.

    python -m floodsim_amd.rocprof_summary trace  <k_kernel_trace.csv> --kernel 'sw_stage<1' --last 1240
    python -m floodsim_amd.rocprof_summary traffic <fetch counter_collection.csv> <write counter_collection.csv> \
        --kernel 'sw_stage<1' --last 1240 --copy-bytes 2147483648
    python -m floodsim_amd.rocprof_summary counters <counter_collection.csv> --last 1240 --clock-ghz 2.4

`traffic` turns FETCH_SIZE and WRITE_SIZE (KiB per dispatch) into HBM bytes per dispatch of the
kernel, each scaled by the ratio of the known size to what the same counter reads for the copy
kernel (fs_copy_kernel, `copy-bytes` read and written per dispatch) in the same run.

`counters` gives, per kernel, the mean of every collected counter over its last N dispatches and,
from SQ_INSTS_VALU, the time its VALU instructions take at the full issue rate of the chip (one
wave64 instruction per 4 cycles on each of 1024 SIMDs at `clock-ghz`, packed FP32 not used).
"""
import argparse
import csv
import json
import re
import statistics
import sys


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _short(name):
    return re.sub(r"\(.*", "", name).replace("void ", "").strip()


def trace_window(path, kernel, last):
    """Per-kernel count / mean / min / max duration (us) and share inside the timed window."""
    rows = [r for r in _rows(path) if r.get("Kind", "KERNEL_DISPATCH") == "KERNEL_DISPATCH"]
    for r in rows:
        r["s"], r["e"] = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
    rows.sort(key=lambda r: r["s"])
    dom = [r for r in rows if re.search(kernel, r["Kernel_Name"])]
    if len(dom) < last:
        raise SystemExit(f"only {len(dom)} dispatches match {kernel!r}")
    win = dom[-last:]
    t0, t1 = win[0]["s"], win[-1]["e"]
    # the window runs from the first timed launch of the dominant kernel to its last; the steps'
    # other kernels fall inside it, except those before the first of these launches in its step
    inside = [r for r in rows if r["s"] >= t0 and r["e"] <= t1]
    by = {}
    for r in inside:
        by.setdefault(_short(r["Kernel_Name"]), []).append((r["e"] - r["s"]) / 1e3)
    total = sum(sum(v) for v in by.values())
    out = []
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        out.append(dict(kernel=k, calls=len(v), mean_us=statistics.fmean(v), min_us=min(v), max_us=max(v),
                        total_ms=sum(v) / 1e3, share=sum(v) / total))
    return dict(window_ms=(t1 - t0) / 1e6, kernels=out)


def traffic(fetch_path, write_path, kernel, last, copy_bytes):
    """HBM bytes per dispatch of `kernel` over its last `last` dispatches, calibrated on the copy kernel."""
    def per(path, counter):
        rows = [r for r in _rows(path) if r["Counter_Name"] == counter]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        k = [float(r["Counter_Value"]) for r in rows if re.search(kernel, r["Kernel_Name"])][-last:]
        c = [float(r["Counter_Value"]) for r in rows if "fs_copy_kernel" in r["Kernel_Name"]]
        if len(k) < last or not c:
            raise SystemExit(f"{path}: {len(k)} dispatches of {kernel!r}, {len(c)} of the copy kernel")
        return statistics.fmean(k), statistics.fmean(c)
    fk, fc = per(fetch_path, "FETCH_SIZE")
    wk, wc = per(write_path, "WRITE_SIZE")
    kib = copy_bytes / 1024.0
    sf, sw = kib / fc, kib / wc
    read_b, write_b = fk * sf * 1024.0, wk * sw * 1024.0
    return dict(fetch_kib=fk, write_kib=wk, copy_fetch_kib=fc, copy_write_kib=wc, fetch_scale=sf, write_scale=sw,
                read_bytes=read_b, write_bytes=write_b, bytes_per_launch=read_b + write_b, launches=last)


def counters(path, last, clock_ghz, simds=1024):
    """Mean counter values per dispatch over each kernel's last `last` dispatches."""
    by = {}
    for r in _rows(path):
        by.setdefault(_short(r["Kernel_Name"]), []).append(r)
    out = {}
    for k, rs in by.items():
        disp = sorted({int(r["Dispatch_Id"]) for r in rs})[-last:]
        keep = set(disp)
        tot = {}
        for r in rs:
            if int(r["Dispatch_Id"]) in keep:
                tot[r["Counter_Name"]] = tot.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        per = {c: v / len(disp) for c, v in tot.items()}
        if "SQ_INSTS_VALU" in per:
            per["valu_issue_ms"] = per["SQ_INSTS_VALU"] * 4.0 / (simds * clock_ghz * 1e9) * 1e3
        out[k] = dict(dispatches=len(disp), per_dispatch=per)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("trace")
    a.add_argument("path")
    a.add_argument("--kernel", required=True)
    a.add_argument("--last", type=int, required=True)
    b = sub.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("--kernel", required=True)
    b.add_argument("--last", type=int, required=True)
    b.add_argument("--copy-bytes", type=float, default=float(2 << 30))
    c = sub.add_parser("counters")
    c.add_argument("path")
    c.add_argument("--last", type=int, required=True)
    c.add_argument("--clock-ghz", type=float, default=2.4)
    args = ap.parse_args(argv)
    if args.cmd == "trace":
        res = trace_window(args.path, args.kernel, args.last)
        print(f"timed window {res['window_ms']:.3f} ms")
        print(f"{'kernel':58s} {'calls':>6s} {'mean us':>10s} {'min us':>9s} {'max us':>9s} {'total ms':>9s} {'share':>6s}")
        for k in res["kernels"]:
            print(f"{k['kernel'][:58]:58s} {k['calls']:6d} {k['mean_us']:10.1f} {k['min_us']:9.1f} {k['max_us']:9.1f} "
                  f"{k['total_ms']:9.2f} {100 * k['share']:5.1f}%")
    elif args.cmd == "counters":
        json.dump(counters(args.path, args.last, args.clock_ghz), sys.stdout, indent=1)
        print()
    else:
        res = traffic(args.fetch, args.write, args.kernel, args.last, args.copy_bytes)
        json.dump(res, sys.stdout, indent=1)
        print()


if __name__ == "__main__":
    main()
