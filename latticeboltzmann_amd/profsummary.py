"""Counter summary of one kernel from rocprofv3 output, as the JSON that
``bench.py`` reads for its roofline (``profiles/roofline_*.json``).

Inputs are the CSV files rocprofv3 writes with ``--output-format csv``: the
``*_kernel_stats.csv`` (or the per-dispatch ``*_kernel_trace.csv``) of a
``--kernel-trace --stats`` run, and the ``*_counter_collection.csv`` of one or
more ``--pmc`` runs of the same command.  For the kernel whose name contains
``--kernel`` the summary keeps the average duration from the trace and the
mean of every counter per dispatch.  ``--grid`` keeps only the dispatches of
that many work-items (one launch shape of a kernel that several variants
share); it needs the per-dispatch trace.  HBM bytes per launch are FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) times 1024, each multiplied by the factor that makes the same
counter read the known byte count of a plain copy kernel (``--calib``: a
counter CSV of ``bench.py --copy-only``, whose relayout kernel reads and
writes ``--calib-bytes`` bytes each).

Usage::

    python -m latticeboltzmann_amd.profsummary --kernel "k_stepk<double, 1, false, 4, 22, 32, 768>" \\
        --key multi20 --trace kt_kernel_stats.csv --pmc p1.csv p2.csv ... \\
        --calib-fetch c1.csv --calib-write c2.csv --calib-bytes 301989888 \\
        --command "..." --out profiles/roofline_f64_cavity_2048.json [--merge]
"""

import argparse
import csv
import json
import os
from collections import defaultdict

CALIB_KERNEL = "k_relayout"


def trace_average_us(path, match, grid=None):
    """(full name, calls, average microseconds) of the kernel matching ``match``,
    from a kernel_stats.csv, or from a kernel_trace.csv (dispatches of ``grid``
    work-items only, when given)."""
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    if rows and "Kernel_Name" in rows[0]:
        durs, name = [], None
        for row in rows:
            if match in row["Kernel_Name"] and (grid is None or _grid(row) == grid):
                name = row["Kernel_Name"]
                durs.append(int(row["End_Timestamp"]) - int(row["Start_Timestamp"]))
        if durs:
            return name, len(durs), sum(durs) / len(durs) / 1e3
    elif grid is not None:
        raise ValueError("a grid filter needs the per-dispatch kernel_trace.csv")
    else:
        for row in rows:
            if match in row["Name"]:
                return row["Name"], int(row["Calls"]), float(row["AverageNs"]) / 1e3
    raise KeyError("no kernel matching %r in %s" % (match, path))


def _grid(row):
    """Work-items of a dispatch (trace CSVs give x, y, z; counter CSVs the total)."""
    if "Grid_Size" in row:
        return int(row["Grid_Size"])
    return int(row["Grid_Size_X"]) * int(row["Grid_Size_Y"]) * int(row["Grid_Size_Z"])


def stats_table(path, top=20):
    """Text table of the ``top`` kernels by total time from a kernel_stats.csv."""
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = ["%-100s %6s %13s %11s %11s %6s" % ("kernel", "calls", "avg_ns", "min_ns", "max_ns", "time%")]
    for r in rows[:top]:
        lines.append("%-100s %6d %13.1f %11d %11d %5.1f%%" % (
            r["Name"][:100], int(r["Calls"]), float(r["AverageNs"]), int(float(r["MinNs"])),
            int(float(r["MaxNs"])), float(r["Percentage"])))
    return "\n".join(lines) + "\n"


def counter_means(paths, match, grid=None):
    """{counter: mean value per dispatch} over the dispatches of the matching
    kernel (of ``grid`` work-items only, when given)."""
    vals = defaultdict(list)
    for path in paths:
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if match in row["Kernel_Name"] and (grid is None or _grid(row) == grid):
                    vals[row["Counter_Name"]].append(float(row["Counter_Value"]))
    if not vals:
        raise KeyError("no dispatch of a kernel matching %r" % match)
    return {k: sum(v) / len(v) for k, v in sorted(vals.items())}


def summarize(kernel, trace, pmc, calib_fetch, calib_write, calib_bytes, grid=None):
    name, calls, avg_us = trace_average_us(trace, kernel, grid)
    c = counter_means(pmc, kernel, grid)
    out = {"kernel": name, "kernel_trace_calls": calls, "kernel_trace_avg_us": round(avg_us, 3),
           "counters_per_dispatch": {k: round(v, 3) for k, v in c.items()}}
    if grid is not None:
        out["grid_work_items"] = grid
    if "SQ_INSTS_VALU" in c:
        out["valu_wave_insts_per_launch"] = c["SQ_INSTS_VALU"]
    if "FETCH_SIZE" in c and "WRITE_SIZE" in c and calib_fetch and calib_write and calib_bytes:
        kib = calib_bytes / 1024.0
        f_scale = kib / counter_means([calib_fetch], CALIB_KERNEL)["FETCH_SIZE"]
        w_scale = kib / counter_means([calib_write], CALIB_KERNEL)["WRITE_SIZE"]
        out["fetch_scale"] = round(f_scale, 4)
        out["write_scale"] = round(w_scale, 4)
        out["hbm_read_bytes_per_launch"] = int(c["FETCH_SIZE"] * 1024 * f_scale)
        out["hbm_write_bytes_per_launch"] = int(c["WRITE_SIZE"] * 1024 * w_scale)
        out["hbm_bytes_per_launch"] = out["hbm_read_bytes_per_launch"] + out["hbm_write_bytes_per_launch"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kernel", required=True, help="substring of the kernel's name")
    ap.add_argument("--key", required=True, help="entry name under per_variant (e.g. multi20)")
    ap.add_argument("--trace", required=True,
                    help="kernel_stats.csv, or kernel_trace.csv (needed with --grid)")
    ap.add_argument("--grid", type=int, default=None,
                    help="only dispatches of this many work-items (one launch shape)")
    ap.add_argument("--pmc", nargs="+", required=True)
    ap.add_argument("--calib-fetch", default=None)
    ap.add_argument("--calib-write", default=None)
    ap.add_argument("--calib-bytes", type=float, default=None)
    ap.add_argument("--command", default="", help="the profiled command, recorded in the file")
    ap.add_argument("--note", default="")
    ap.add_argument("--out", required=True)
    ap.add_argument("--merge", action="store_true", help="add to an existing file's per_variant")
    ap.add_argument("--stats-out", default=None,
                    help="also write the trace's kernel table (text) to this file")
    ap.add_argument("--stats-title", default="")
    ap.add_argument("--stats-csv", default=None,
                    help="kernel_stats.csv for the --stats-out table (default: --trace)")
    a = ap.parse_args(argv)
    if a.stats_out:
        with open(a.stats_out, "w") as fh:
            if a.stats_title:
                fh.write(a.stats_title.rstrip("\n") + "\n")
            if a.command:
                fh.write("command (from the repo root on an MI355X, ROCm 7.2):\n  %s\n" % a.command)
            fh.write("\n" + stats_table(a.stats_csv or a.trace))
    entry = summarize(a.kernel, a.trace, a.pmc, a.calib_fetch, a.calib_write, a.calib_bytes,
                      a.grid)
    if a.command:
        entry["command"] = a.command
    if a.note:
        entry["note"] = a.note
    doc = {"per_variant": {}}
    if a.merge and os.path.exists(a.out):
        with open(a.out) as fh:
            doc = json.load(fh)
    doc.setdefault("per_variant", {})[a.key] = entry
    with open(a.out, "w") as fh:
        json.dump(doc, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print(json.dumps({a.key: {k: entry[k] for k in entry if k != "counters_per_dispatch"}}))


if __name__ == "__main__":
    main()
