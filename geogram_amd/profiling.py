"""Summaries of rocprofv3 output for the committed profiles.

Two commands, each reading the CSV files of one or more rocprofv3 runs and
writing a short markdown summary:

  python -m geogram_amd.profiling trace --csv-dir D --prefix P --kernel K \\
      --title T --command C --out profiles/NAME.md
      kernel statistics (P_kernel_stats.csv, top rows) and the per-launch
      durations and launch shape of the kernels whose name contains K
      (P_kernel_trace.csv);

  python -m geogram_amd.profiling counters --fetch F.csv --write W.csv \\
      [--sq S1.csv ...] --kernel K --title T --command C ... --out profiles/NAME.md
      HBM bytes per launch of kernel K from FETCH_SIZE / WRITE_SIZE (KiB per
      dispatch), scaled by what the same counters read for the largest
      __amd_rocclr_copyBuffer dispatch of the run (a device-to-device copy of
      known size, bench.py --copy-calib), and issue counters per launch.

Both stamp the summary with ``kernel_sources_sha1``, the hash of the kernel
sources it was taken of; bench.py reports a figure from a summary only when
that hash matches the sources it runs.
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import os
import statistics
import sys

_CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
KERNEL_SOURCES = ("clip_cell.h", "wave.h", "types.h", "common.h", "exact_det.h", "backend_hip.hip",
                  "ot_kernels.h", "surface.h", "mlp.h")

# MI355X: 256 CUs x 4 SIMDs; SQ_BUSY_CYCLES is summed over 32 shader engines
N_SIMDS = 1024
N_SE = 32


def kernel_sources_sha1() -> str:
    """Hash of the gfx950 kernel sources (the profiled code)."""
    h = hashlib.sha1()
    for f in KERNEL_SOURCES:
        with open(os.path.join(_CSRC, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def read_figure(path: str, key: str):
    """The `key: value` line of a committed summary, or None unless the
    summary was taken of the current kernel sources."""
    vals = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" in line and not line.startswith(("|", "#", " ", "`")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
    except OSError:
        return None
    if vals.get("kernel_sources_sha1") != kernel_sources_sha1() or key not in vals:
        return None
    try:
        return float(vals[key])
    except ValueError:
        return None


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _col(row, *names):
    for n in names:
        if n in row:
            return row[n]
    raise KeyError(f"none of {names} in {list(row)}")


def _short(name, width=80):
    return name if len(name) <= width else name[: width - 3] + "..."


def trace_summary(csv_dir, prefix, kernel, title, command, notes=""):
    stats = _rows(os.path.join(csv_dir, f"{prefix}_kernel_stats.csv"))
    trace = _rows(os.path.join(csv_dir, f"{prefix}_kernel_trace.csv"))
    out = [f"# {title}", "", notes, "", "Command:", "", "```", command, "```", "",
           f"kernel_sources_sha1: {kernel_sources_sha1()}", "",
           f"## Kernel statistics (`{prefix}_kernel_stats.csv`, top rows)", "",
           "| kernel | calls | total ms | average ms | % |", "|---|---:|---:|---:|---:|"]
    for r in stats[:14]:
        out.append(f"| `{_short(_col(r, 'Name', 'KernelName'))}` | {_col(r, 'Calls')} | "
                   f"{float(_col(r, 'TotalDurationNs')) / 1e6:.3f} | "
                   f"{float(_col(r, 'AverageNs')) / 1e6:.3f} | {float(_col(r, 'Percentage')):.2f} |")
    for kname in sorted({_col(r, "Kernel_Name") for r in trace if kernel in _col(r, "Kernel_Name")}):
        launches = [r for r in trace if _col(r, "Kernel_Name") == kname]
        durs = [(int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e6
                for r in launches]
        r0 = launches[-1]
        out += ["", f"## Launches of `{_short(kname, 120)}`", "",
                "| launch | ms |", "|---:|---:|"]
        out += [f"| {i} | {d:.4f} |" for i, d in enumerate(durs[:40])]
        if len(durs) > 40:
            out.append(f"| ... | ({len(durs)} launches) |")
        out += ["", f"launches: {len(durs)}", f"mean_ms: {statistics.fmean(durs):.5f}",
                f"median_ms: {statistics.median(durs):.5f}", "",
                "Launch shape: grid {}, work-group {}, LDS {} B, VGPR {}, SGPR {}, scratch {} "
                "B/lane.".format(_col(r0, "Grid_Size", "Grid_Size_X"),
                                 _col(r0, "Workgroup_Size", "Workgroup_Size_X"),
                                 _col(r0, "LDS_Block_Size", "Lds_Size"),
                                 _col(r0, "VGPR_Count", "Arch_VGPR_Count"),
                                 _col(r0, "SGPR_Count"), _col(r0, "Scratch_Size"))]
    return "\n".join(out) + "\n"


def _per_dispatch(rows, kernel):
    """{dispatch id: {counter: value}} of the kernels whose name contains
    kernel (values of one counter in one dispatch summed over its
    instances)."""
    out = {}
    for r in rows:
        if kernel not in _col(r, "Kernel_Name"):
            continue
        d = int(_col(r, "Dispatch_Id", "Correlation_Id"))
        c = _col(r, "Counter_Name")
        out.setdefault(d, {})
        out[d][c] = out[d].get(c, 0.0) + float(_col(r, "Counter_Value"))
    return dict(sorted(out.items()))


def counters_summary(fetch_csv, write_csv, sq_csvs, kernel, title, commands, skip_first,
                     calib_bytes, notes="", sq_skip_first=None):
    out = [f"# {title}", "", notes, "", "Commands (one counter set per run):", "", "```"]
    out += commands + ["```", "", f"kernel_sources_sha1: {kernel_sources_sha1()}", ""]
    scale = {}
    traffic = 0.0
    for label, path, cname in (("FETCH_SIZE", fetch_csv, "FETCH_SIZE"),
                               ("WRITE_SIZE", write_csv, "WRITE_SIZE")):
        rows = _rows(path)
        cal = _per_dispatch(rows, "__amd_rocclr_copyBuffer")
        ref = max((v.get(cname, 0.0) for v in cal.values()), default=0.0)
        scale[label] = (calib_bytes / 1024.0) / ref if ref > 0 else float("nan")
        k = _per_dispatch(rows, kernel)
        vals = [v.get(cname, 0.0) for v in k.values()][skip_first:]
        mean_kib = statistics.fmean(vals) if vals else float("nan")
        b = mean_kib * 1024.0 * scale[label]
        traffic += b
        out += [f"{label}: {len(vals)} launches, mean {mean_kib:,.1f} KiB raw; copy "
                f"calibration reads {ref:,.1f} KiB for {calib_bytes / 1024:,.0f} KiB moved, "
                f"scale {scale[label]:.4f}: {b:,.0f} bytes per launch", ""]
    out += [f"traffic_bytes_per_launch: {traffic:.0f}", ""]
    sq = {}
    for path in sq_csvs:
        for d, v in _per_dispatch(_rows(path), kernel).items():
            for c, x in v.items():
                sq.setdefault(c, []).append(x)
    if sq:
        k = skip_first if sq_skip_first is None else sq_skip_first
        sq = {c: (v[k:] or v) for c, v in sq.items()}
        out += ["| counter | mean per launch |", "|---|---:|"]
        for c in sorted(sq):
            out.append(f"| {c} | {statistics.fmean(sq[c]):,.1f} |")
        if "SQ_ACTIVE_INST_VALU" in sq and "SQ_BUSY_CYCLES" in sq:
            v = statistics.fmean(sq["SQ_ACTIVE_INST_VALU"])
            busy = statistics.fmean(sq["SQ_BUSY_CYCLES"]) / N_SE
            out += ["", f"valu_busy_frac: {4.0 * v / (N_SIMDS * busy):.4f}"]
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m geogram_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("--csv-dir", required=True)
    t.add_argument("--prefix", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--command", required=True)
    t.add_argument("--notes", default="")
    t.add_argument("--out", required=True)
    c = sub.add_parser("counters")
    c.add_argument("--fetch", required=True)
    c.add_argument("--write", required=True)
    c.add_argument("--sq", nargs="*", default=[])
    c.add_argument("--kernel", required=True)
    c.add_argument("--title", required=True)
    c.add_argument("--command", action="append", required=True)
    c.add_argument("--skip-first", type=int, default=0,
                   help="launches to leave out (work-counter pass, warmup)")
    c.add_argument("--sq-skip-first", type=int, default=None,
                   help="launches of the SQ runs to leave out (default: --skip-first)")
    c.add_argument("--calib-bytes", type=float, default=float(256 << 20))
    c.add_argument("--notes", default="")
    c.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        text = trace_summary(a.csv_dir, a.prefix, a.kernel, a.title, a.command, a.notes)
    else:
        text = counters_summary(a.fetch, a.write, a.sq, a.kernel, a.title, a.command, a.skip_first,
                                a.calib_bytes, a.notes, a.sq_skip_first)
    with open(a.out, "w") as f:
        f.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
