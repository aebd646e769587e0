"""Busy fraction and inter-kernel gaps of a rocprofv3 kernel trace.

    python -m fpc_amd.trace_gaps TRACE_DIR [LAST_MS]

reads the *kernel_trace.csv under TRACE_DIR (rocprofv3 --kernel-trace
--output-format csv) and reports, over the whole trace or its last LAST_MS
milliseconds: the number of kernels, the span, the time covered by at least one
kernel, and the idle time between kernels binned by gap length.  Gaps of a few
microseconds are launch-pipeline gaps (what a captured graph could remove);
longer ones are host waits."""
import collections
import csv
import glob
import sys

BINS = ((2e3, "<2us"), (5e3, "<5us"), (1e4, "<10us"), (5e4, "<50us"), (float("inf"), ">=50us"))


def load(trace_dir):
    files = glob.glob(trace_dir.rstrip("/") + "/**/*kernel_trace.csv", recursive=True)
    if not files:
        raise FileNotFoundError(f"no *kernel_trace.csv under {trace_dir}")
    rows = []
    for f in files:
        rows += [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in csv.DictReader(open(f))]
    return sorted(rows)


def analyse(rows, last_ns=None):
    if last_ns is not None:
        t_end = max(e for _, e, _ in rows)
        rows = [r for r in rows if r[0] >= t_end - last_ns]
    span = max(e for _, e, _ in rows) - rows[0][0]
    busy, end, gaps = 0, rows[0][0], []
    for s, e, _ in rows:
        if s > end:
            gaps.append(s - end)
        busy += max(0, e - max(s, end))
        end = max(end, e)
    hist = collections.OrderedDict((name, 0) for _, name in BINS)
    for g in gaps:
        for lim, name in BINS:
            if g < lim:
                hist[name] += g
                break
    return {"kernels": len(rows), "span_ms": span / 1e6, "busy_ms": busy / 1e6, "busy_frac": busy / span,
            "idle_ms": sum(gaps) / 1e6, "gaps": len(gaps), "idle_ms_by_gap": {k: v / 1e6 for k, v in hist.items()},
            "largest_gaps_us": [g / 1e3 for g in sorted(gaps)[-5:]]}


def main(argv):
    rows = load(argv[1])
    last = float(argv[2]) * 1e6 if len(argv) > 2 else None
    r = analyse(rows, last)
    print(f"kernels {r['kernels']}, span {r['span_ms']:.2f} ms, busy {r['busy_ms']:.2f} ms "
          f"({100 * r['busy_frac']:.1f}%), idle {r['idle_ms']:.2f} ms in {r['gaps']} gaps")
    for k, v in r["idle_ms_by_gap"].items():
        print(f"  idle in gaps {k:7s}: {v:8.3f} ms")
    print("  largest gaps (us):", [round(g, 1) for g in r["largest_gaps_us"]])


if __name__ == "__main__":
    main(sys.argv)
