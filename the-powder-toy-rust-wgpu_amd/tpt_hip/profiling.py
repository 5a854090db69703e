"""Summaries of rocprofv3 runs of bench.py (rocpd SQLite databases).

    python -m tpt_hip.profiling --trace T.db --fetch F.db --write W.db \
        --copy-fetch CF.db --copy-write CW.db --copy-bytes 268435456 \
        --commands CMDS.txt --out profiles/tick_2048x2048

writes `<out>.md` (per-kernel durations from the kernel trace, FETCH_SIZE and
WRITE_SIZE per dispatch, calibrated HBM bytes) and `<out>_traffic.json` (the
calibrated HBM bytes per dispatch by kernel, which bench.py reports as the
roofline's `traffic`).  With only `--trace` (and `--commands`, `--out`) it
writes the kernel-time table alone, `--match` keeping the kernels whose name
contains one of the given substrings.

Calibration: the same counters are collected for a device-to-device copy of
`--copy-bytes` bytes (bench.py --calibrate-copy); each counter's scale is the
copy's true bytes moved (read once, written once) over what the counter reads
for it.
"""
import argparse
import collections
import json
import re
import sqlite3


def short(name):
    m = re.search(r"::(\w+)\(", name) or re.search(r"(\w+)\(", name)
    return m.group(1) if m else name[:60]


def kernel_stats(db):
    """[(kernel, calls, total_us, avg_us, percent)] from the kernel trace."""
    con = sqlite3.connect(db)
    rows = []
    for name, calls, total, avg, pct in con.execute("select * from top_kernels"):
        rows.append((short(name), int(calls), float(total), float(avg), float(pct)))
    return rows


def counter_means(db, counter):
    """{kernel: (dispatches, mean value per dispatch)} of one PMC counter."""
    con = sqlite3.connect(db)
    agg = collections.defaultdict(list)
    for name, val in con.execute(
            "select kernel_name, value from counters_collection where counter_name = ?", (counter,)):
        agg[short(name)].append(float(val))
    return {k: (len(v), sum(v) / len(v)) for k, v in agg.items()}


def copy_scale(db, counter, copy_bytes):
    """Bytes per counter unit: the run holds the copy and little else, so its
    largest per-dispatch value is the copy's."""
    vals = counter_means(db, counter)
    best = max(vals.values(), key=lambda v: v[1])  # the copy dominates its own run
    return copy_bytes / best[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--copy-bytes", type=int)
    ap.add_argument("--commands", required=True, help="text file with the rocprofv3 commands that were run")
    ap.add_argument("--title", default="Whole tick, 2048x2048 canvas")
    ap.add_argument("--match", nargs="*", default=None, help="keep kernels whose name contains one of these")
    ap.add_argument("--mfma", help="rocprofv3 database of a --pmc MfmaUtil MfmaFlopsBF16 run: adds MFMA "
                                   "utilisation and bf16 MFMA FLOP per dispatch to the kernel-time table")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()

    trace = kernel_stats(a.trace)
    if a.match:
        trace = [t for t in trace if any(m in t[0] for m in a.match)]
    if not a.fetch:
        lines = [f"# {a.title}", "", "Commands (run from /tmp with TMPDIR=/tmp):", "", "```",
                 open(a.commands).read().strip(), "```", ""]
        if a.mfma:
            util = counter_means(a.mfma, "MfmaUtil")
            flops = counter_means(a.mfma, "MfmaFlopsBF16")
            lines += ["MfmaUtil: SQ_VALU_MFMA_BUSY_CYCLES over GRBM_GUI_ACTIVE x SIMDs, in percent, per dispatch; "
                      "MfmaFlopsBF16: bf16 MFMA FLOP per dispatch (SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512); "
                      "TFLOP/s = that over the kernel's average duration.", "",
                      "| kernel | calls | avg us | total us | % GPU time | MFMA util % | bf16 MFMA GFLOP/dispatch "
                      "| TFLOP/s |", "|---|---|---|---|---|---|---|---|"]
            for name, calls, total, avg, pct in trace:
                u = util.get(name, (0, 0.0))[1]
                f = flops.get(name, (0, 0.0))[1]
                tf = f / (avg * 1e-6) / 1e12 if avg > 0 else 0.0
                lines.append(f"| {name} | {calls} | {avg:.1f} | {total:.0f} | {pct:.1f} | {u:.1f} | {f / 1e9:.2f} "
                             f"| {tf:.0f} |")
        else:
            lines += ["| kernel | calls | avg us | total us | % GPU time |", "|---|---|---|---|---|"]
            for name, calls, total, avg, pct in trace:
                lines.append(f"| {name} | {calls} | {avg:.1f} | {total:.0f} | {pct:.1f} |")
        with open(a.out + ".md", "w") as fh:
            fh.write("\n".join(lines) + "\n")
        return
    fetch = counter_means(a.fetch, "FETCH_SIZE")
    write = counter_means(a.write, "WRITE_SIZE")
    fscale = copy_scale(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    wscale = copy_scale(a.copy_write, "WRITE_SIZE", a.copy_bytes)

    traffic = {}
    lines = [f"# {a.title}", "", "Commands (run from /tmp with TMPDIR=/tmp):", "", "```",
             open(a.commands).read().strip(), "```", "",
             f"HBM calibration: a {a.copy_bytes >> 20} MiB device-to-device copy moves {a.copy_bytes} bytes "
             f"each way; FETCH_SIZE x {fscale:.1f} B and WRITE_SIZE x {wscale:.1f} B reproduce it "
             f"(counters in KiB: x {fscale / 1024:.3f} and x {wscale / 1024:.3f}).", "",
             "| kernel | calls | avg us | % GPU time | FETCH_SIZE/dispatch | WRITE_SIZE/dispatch | HBM MB/dispatch |",
             "|---|---|---|---|---|---|---|"]
    for name, calls, total, avg, pct in trace:
        f = fetch.get(name, (0, 0.0))[1]
        w = write.get(name, (0, 0.0))[1]
        hbm = f * fscale + w * wscale
        traffic[name] = {"fetch": f, "write": w, "hbm_bytes": round(hbm)}
        lines.append(f"| {name} | {calls} | {avg:.1f} | {pct:.1f} | {f:.0f} | {w:.0f} | {hbm / 1e6:.1f} |")
    with open(a.out + ".md", "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(a.out + "_traffic.json", "w") as fh:
        json.dump({"fetch_scale_bytes": fscale, "write_scale_bytes": wscale, "kernels": traffic}, fh, indent=1)


if __name__ == "__main__":
    main()
