"""rocprofv3 summaries for the benchmark's roofline figures.

Reads the CSV files that `rocprofv3 --kernel-trace --stats --output-format csv`
and `rocprofv3 --pmc FETCH_SIZE|WRITE_SIZE --output-format csv` write, and
produces the short text summaries kept under profiles/ plus the HBM traffic
per launch of the dominant kernel:

  traffic = FETCH_SIZE * 1024 * kf + WRITE_SIZE * 1024 * kw   [bytes/launch]

where FETCH_SIZE and WRITE_SIZE are the counters' per-dispatch values (KiB)
and kf, kw scale them so that a plain copy kernel of known size reads its
known byte counts (the copy kernel of fdfault_amd/csrc/copy_bw.hip moving
`copy_mib` MiB each way per launch).

Usage (after the rocprofv3 runs):
  python -m fdfault_amd.profiling stats  <kernel_stats.csv> <out.md> <title> <command>
  python -m fdfault_amd.profiling line   <bench.log> <out.md> <title> <command>   (append a bench line)
  python -m fdfault_amd.profiling stages <kernel_trace.csv> <kernel-substring> <launches-per-stage> <out.md>
  python -m fdfault_amd.profiling traffic <key> <kernel-substring> <fetch.csv> <write.csv>
         <copy_fetch.csv> <copy_write.csv> <copy_mib> <points_per_launch> <out.md> <command...>
The traffic mode also records the result under <key> in profiles/traffic.json.
"""
from __future__ import annotations

import csv
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic.json")


def kernel_stats(path):
    """Rows of a rocprofv3 kernel_stats.csv: (name, calls, total_ns, avg_ns, percent)."""
    out = []
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            out.append((row["Name"], int(row["Calls"]), float(row["TotalDurationNs"]), float(row["AverageNs"]),
                        float(row["Percentage"])))
    out.sort(key=lambda r: -r[2])
    return out


def counter_per_dispatch(path, kernel_sub, counter):
    """Mean per-dispatch value of `counter` over dispatches whose kernel name
    contains `kernel_sub` (a counter CSV holds one row per dispatch and counter)."""
    vals = []
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            if row["Counter_Name"] == counter and kernel_sub in row["Kernel_Name"]:
                vals.append(float(row["Counter_Value"]))
    if not vals:
        raise ValueError("no %s rows for kernel %r in %s" % (counter, kernel_sub, path))
    return sum(vals) / len(vals), len(vals)


def write_stats(csv_path, out_md, title, command):
    rows = kernel_stats(csv_path)
    with open(out_md, "w") as fh:
        fh.write("# %s\n\n" % title)
        fh.write("Command:\n\n```\n%s\n```\n\n" % command)
        fh.write("| kernel | calls | avg (us) | total (ms) | % |\n|---|---|---|---|---|\n")
        for name, calls, tot, avg, pct in rows[:12]:
            short = name if len(name) < 90 else name[:87] + "..."
            fh.write("| `%s` | %d | %.1f | %.3f | %.1f |\n" % (short.replace("|", "\\|"), calls, avg / 1e3, tot / 1e6, pct))


def write_traffic(key, kernel_sub, fetch_csv, write_csv, cfetch_csv, cwrite_csv, copy_mib, points, out_md, command):
    copy_bytes = copy_mib * 1024 * 1024
    cf, _ = counter_per_dispatch(cfetch_csv, "copy_kernel", "FETCH_SIZE")
    cw, _ = counter_per_dispatch(cwrite_csv, "copy_kernel", "WRITE_SIZE")
    kf, kw = copy_bytes / (cf * 1024.), copy_bytes / (cw * 1024.)
    f, nf = counter_per_dispatch(fetch_csv, kernel_sub, "FETCH_SIZE")
    w, nw = counter_per_dispatch(write_csv, kernel_sub, "WRITE_SIZE")
    fb, wb = f * 1024. * kf, w * 1024. * kw
    rec = {"kernel": kernel_sub, "fetch_bytes": round(fb), "write_bytes": round(wb), "traffic": round(fb + wb),
           "bytes_per_point": round((fb + wb) / points, 2), "points_per_launch": points,
           "fetch_scale": round(kf, 4), "write_scale": round(kw, 4), "summary": os.path.relpath(out_md, ROOT)}
    with open(out_md, "w") as fh:
        fh.write("# HBM traffic of `%s` (%s)\n\n" % (kernel_sub, key))
        fh.write("Commands:\n\n```\n%s\n```\n\n" % command)
        fh.write("| quantity | value |\n|---|---|\n")
        fh.write("| copy kernel bytes each way per launch | %d |\n" % copy_bytes)
        fh.write("| copy FETCH_SIZE / WRITE_SIZE (KiB per dispatch) | %.1f / %.1f |\n" % (cf, cw))
        fh.write("| scale factors kf / kw | %.4f / %.4f |\n" % (kf, kw))
        fh.write("| kernel FETCH_SIZE (KiB per dispatch, %d dispatches) | %.1f |\n" % (nf, f))
        fh.write("| kernel WRITE_SIZE (KiB per dispatch, %d dispatches) | %.1f |\n" % (nw, w))
        fh.write("| HBM read bytes per launch | %d |\n" % fb)
        fh.write("| HBM write bytes per launch | %d |\n" % wb)
        fh.write("| points per launch | %d |\n" % points)
        fh.write("| HBM bytes per point (read + write) | %.1f + %.1f = %.1f |\n" % (fb / points, wb / points, (fb + wb) / points))
    data = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as fh:
            data = json.load(fh)
    data[key] = rec
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    return rec


def append_stage_times(trace_csv, kernel_sub, per_stage, out_md):
    """Append the kernel's launch durations grouped by RK stage (from a
    rocprofv3 kernel_trace.csv, launches in order, `per_stage` launches per
    stage, stages of a step in order) to a summary: the first and last
    stages of a step move less df traffic than the middle ones."""
    d = []
    with open(trace_csv, newline="") as fh:
        for row in csv.DictReader(fh):
            if kernel_sub in row["Kernel_Name"]:
                d.append((int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3)
    nst = 5
    per_step = nst * per_stage
    steps = len(d) // per_step
    with open(out_md, "a") as fh:
        fh.write("\nLaunch durations by RK stage (us, mean over %d steps x %d launches per stage):\n\n" % (steps, per_stage))
        fh.write("| stage | df traffic | mean (us) |\n|---|---|---|\n")
        for st in range(nst):
            v = [d[k * per_step + st * per_stage + b] for k in range(steps) for b in range(per_stage)]
            what = "no df load (A = 0)" if st == 0 else ("no df store" if st == nst - 1 else "df load + store")
            fh.write("| %d | %s | %.1f |\n" % (st, what, sum(v) / len(v)))


def _flat(d, pre=""):
    for k, v in d.items():
        if isinstance(v, dict):
            yield from _flat(v, pre + k + ".")
        else:
            yield pre + k, v


def append_bench_line(log, out_md, title, command):
    """Append one bench.py JSON line (the last line of `log` that parses) to a
    summary as a section: the command and every field, nested ones flattened
    as a.b."""
    line = None
    with open(log) as fh:
        for ln in fh:
            ln = ln.strip()
            if ln.startswith("{"):
                try:
                    line = json.loads(ln)
                except ValueError:
                    pass
    if line is None:
        raise ValueError("no JSON line in %s" % log)
    with open(out_md, "a") as fh:
        fh.write("\n## %s\n\n```\n%s\n```\n\n| field | value |\n|---|---|\n" % (title, command))
        for k, v in _flat(line):
            txt = ("%.6g" % v) if isinstance(v, float) else (json.dumps(v) if isinstance(v, (list, str)) else str(v))
            fh.write("| %s | %s |\n" % (k, txt))


def lookup_traffic(key):
    """Recorded traffic for a workload key, or None."""
    if not os.path.exists(TRAFFIC_JSON):
        return None
    with open(TRAFFIC_JSON) as fh:
        return json.load(fh).get(key)


if __name__ == "__main__":
    a = sys.argv[1:]
    if a and a[0] == "stats" and len(a) == 5:
        write_stats(a[1], a[2], a[3], a[4])
    elif a and a[0] == "line" and len(a) == 5:
        append_bench_line(a[1], a[2], a[3], a[4])
    elif a and a[0] == "stages" and len(a) == 5:
        append_stage_times(a[1], a[2], int(a[3]), a[4])
    elif a and a[0] == "traffic" and len(a) >= 11:
        print(json.dumps(write_traffic(a[1], a[2], a[3], a[4], a[5], a[6], float(a[7]), int(a[8]), a[9],
                                       " ".join(a[10:]))))
    else:
        print(__doc__)
        sys.exit(2)
