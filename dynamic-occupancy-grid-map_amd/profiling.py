"""Summaries of rocprofv3 outputs for this library (kernel-trace statistics and counter passes).

    python dynamic-occupancy-grid-map_amd/profiling.py trace <run_results.db> <title> <command>
    python dynamic-occupancy-grid-map_amd/profiling.py pmc <dir> <calibration MiB> <title> <command>

`pmc` expects the databases of four counter passes under <dir>: fetch/ (FETCH_SIZE), write/ (WRITE_SIZE),
sq/ and sq2/ (SQ_* counters), each `<pass>/run_results.db`, from runs that also dispatched one
`k_copy_bytes` kernel copying <calibration MiB> (bench.py --copy-calibration-mb).  FETCH_SIZE and
WRITE_SIZE are scaled by what they read for that copy of known size; the traffic table lists HBM bytes
per launch next to the kernel's duration in the same run.
"""
from __future__ import annotations

import collections
import csv
import re
import sqlite3
import sys


def short_name(name: str) -> str:
    s = re.sub(r"\(.*", "", name)
    return s.replace("dogm_hip::", "").replace("void ", "").strip()


def _trace_rows(path: str):
    """(name, grid, calls, total ns, average ns) per kernel and grid size from a kernel-trace database
    (grid None for the kernel_stats.csv of `--output-format csv`)."""
    if path.endswith(".csv"):
        return [(r["Name"], None, int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]))
                for r in csv.DictReader(open(path))]
    c = sqlite3.connect(path)
    return [(n, int(g), int(k), float(t), float(t) / int(k)) for n, g, k, t in
            c.execute("select name, grid_x, count(*), sum(duration) from kernels group by name, grid_x")]


def trace_summary(trace: str, title: str, command: str) -> str:
    rows = _trace_rows(trace)
    total = sum(r[3] for r in rows)
    out = [f"# {title}", "", "Command (from the repository root, after `export TMPDIR=/tmp`):", "", f"    {command}", "",
           "All dispatches of the run (warm-up and timed steps alike), per kernel and grid size (work-items; a "
           "launch's grid covers the vehicles of its step or lane).", "",
           "| kernel | grid | calls | total us | avg us | % |", "|---|---:|---:|---:|---:|---:|"]
    for name, grid, calls, t, avg in sorted(rows, key=lambda r: -r[3]):
        out.append(f"| `{short_name(name)}` | {grid if grid is not None else ''} | {calls} | {t / 1e3:.1f} | "
                   f"{avg / 1e3:.2f} | {100 * t / total:.2f} |")
    return "\n".join(out) + "\n"


def _load(path: str):
    c = sqlite3.connect(path)
    rows = c.execute("select dispatch_id, kernel_name, grid_size, counter_name, value, duration "
                     "from counters_collection").fetchall()
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = collections.defaultdict(dict)
    for disp, name, grid, cn, val, dt in rows:
        key = (short_name(name), int(grid))
        vals[key][cn].append(float(val))
        dur[key][disp] = float(dt)
    return vals, dur


def _avg(x):
    return sum(x) / len(x) if x else 0.0


def pmc_summary(d: str, cal_mib: int, title: str, command: str) -> str:
    fe, dur = _load(f"{d}/fetch/run_results.db")
    wr, _ = _load(f"{d}/write/run_results.db")
    sq, _ = _load(f"{d}/sq/run_results.db")
    sq2, _ = _load(f"{d}/sq2/run_results.db")
    cal_kib = cal_mib * 1024.0
    cf = [v for (k, g), cs in fe.items() if k.startswith("k_copy_bytes") for v in cs["FETCH_SIZE"]]
    cw = [v for (k, g), cs in wr.items() if k.startswith("k_copy_bytes") for v in cs["WRITE_SIZE"]]
    ff, fw = cal_kib / _avg(cf), cal_kib / _avg(cw)
    out = [f"# {title}", "", "Counter passes (one counter group per run):", "", f"    {command}", "",
           f"Calibration: one `k_copy_bytes` dispatch copying {cal_mib} MiB reads FETCH_SIZE {_avg(cf):.0f} KiB and "
           f"WRITE_SIZE {_avg(cw):.0f} KiB, so FETCH_SIZE is scaled by {ff:.3f} and WRITE_SIZE by {fw:.3f}.", "",
           "Per launch (grid = work-items; batched launches cover every vehicle of the step). "
           "`valu/wave`, `lds/wave` = SQ_INSTS_VALU, SQ_INSTS_LDS per wave; `bank conflicts/wave` = "
           "SQ_LDS_BANK_CONFLICT per wave.", "",
           "| kernel | grid | calls | avg us | HBM read MB | HBM write MB | HBM GB/s | valu/wave | lds/wave | bank conflicts/wave |",
           "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    keys = sorted(fe, key=lambda k: -_avg(list(dur[k].values())) * len(dur[k]))
    for k in keys:
        name, grid = k
        if name.startswith("__amd") or name.startswith("k_copy_bytes"):
            continue
        t_us = _avg(list(dur[k].values())) / 1e3
        rd = _avg(fe[k]["FETCH_SIZE"]) * ff * 1024
        wb = _avg(wr[k]["WRITE_SIZE"]) * fw * 1024
        s, s2 = sq.get(k, {}), sq2.get(k, {})
        w = _avg(s.get("SQ_WAVES", [])) or 1.0
        out.append(f"| `{name}` | {grid} | {len(dur[k])} | {t_us:.1f} | {rd / 1e6:.1f} | {wb / 1e6:.1f} | "
                   f"{(rd + wb) / (t_us * 1e-6) / 1e9:.0f} | {_avg(s.get('SQ_INSTS_VALU', [])) / w:.0f} | "
                   f"{_avg(s.get('SQ_INSTS_LDS', [])) / w:.0f} | {_avg(s2.get('SQ_LDS_BANK_CONFLICT', [])) / w:.1f} |")
    out += ["", "Machine-readable traffic per launch (HBM bytes read + written, scaled), used by bench.py:", ""]
    for k in keys:
        name, grid = k
        if name.startswith("__amd") or name.startswith("k_copy_bytes"):
            continue
        rd = _avg(fe[k]["FETCH_SIZE"]) * ff * 1024
        wb = _avg(wr[k]["WRITE_SIZE"]) * fw * 1024
        out.append(f"    traffic {name.replace(' ', '')} grid={grid} calls={len(dur[k])} bytes={int(rd + wb)}")
    return "\n".join(out) + "\n"


def lookup_traffic(summary_md: str, kernel_prefix: str) -> int | None:
    """HBM bytes per launch recorded for a kernel: entries whose name starts with `kernel_prefix` at
    the largest grid recorded for it (the batched launches), averaged over template instances
    weighted by their launches."""
    try:
        lines = open(summary_md).read().splitlines()
    except OSError:
        return None
    ent = []
    for ln in lines:
        m = re.match(r"\s+traffic (\S+) grid=(\d+) calls=(\d+) bytes=(\d+)", ln)
        if m and (m.group(1) == kernel_prefix or m.group(1).startswith(kernel_prefix + "<")):
            ent.append((int(m.group(2)), int(m.group(3)), int(m.group(4))))
    if not ent:
        return None
    gmax = max(e[0] for e in ent)
    sel = [e for e in ent if e[0] == gmax]
    calls = sum(e[1] for e in sel)
    return int(sum(e[1] * e[2] for e in sel) / calls)


def lookup_valu(summary_md: str, kernel_name: str) -> tuple[int, float] | None:
    """(grid, VALU instructions per wave) of a kernel's row at the largest grid in a counter summary
    table (kernel | grid | calls | avg us | read MB | write MB | GB/s | valu/wave | ...); a name
    without template arguments also matches its template instances."""
    try:
        lines = open(summary_md).read().splitlines()
    except OSError:
        return None
    best = None
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        name = cells[0].strip("`") if cells else ""
        if len(cells) < 8 or not (name == kernel_name or name.startswith(kernel_name + "<")):
            continue
        try:
            grid, valu = int(cells[1]), float(cells[7])
        except ValueError:
            continue
        if best is None or grid > best[0]:
            best = (grid, valu)
    return best


def marker_summary(csv_path: str, title: str, command: str) -> str:
    """roctx ranges of a run (rocprofv3 --marker-trace --stats --output-format csv: the marker statistics
    CSV with Name, Calls, TotalDurationNs, AverageNs): host time spent inside each stage's range, i.e.
    enqueueing its kernels (the kernels themselves run asynchronously on the device)."""
    rows = [(r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]))
            for r in csv.DictReader(open(csv_path))]
    out = [f"# {title}", "", "Command (from the repository root, after `export TMPDIR=/tmp`):", "", f"    {command}", "",
           "Host time inside each roctx range of the library (a stage's range covers the enqueue of its kernels; the "
           "kernels run asynchronously, so device time is in the kernel trace).", "",
           "| range | calls | total us | avg us |", "|---|---:|---:|---:|"]
    for name, calls, t, avg in sorted(rows, key=lambda r: -r[2]):
        out.append(f"| `{name}` | {calls} | {t / 1e3:.1f} | {avg / 1e3:.2f} |")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    if len(sys.argv) >= 5 and sys.argv[1] == "trace":
        print(trace_summary(sys.argv[2], sys.argv[3], sys.argv[4]), end="")
    elif len(sys.argv) >= 5 and sys.argv[1] == "markers":
        print(marker_summary(sys.argv[2], sys.argv[3], sys.argv[4]), end="")
    elif len(sys.argv) >= 6 and sys.argv[1] == "pmc":
        print(pmc_summary(sys.argv[2], int(sys.argv[3]), sys.argv[4], sys.argv[5]), end="")
    else:
        print(__doc__)
        sys.exit(2)
