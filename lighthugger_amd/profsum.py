"""Summaries of rocprofv3 output (kernel-trace statistics and HBM counters).

rocprofv3 writes its results as SQLite databases (``*_results.db``) or, with
``--output-format csv``, as ``kernel_stats.csv`` / ``counter_collection.csv``.
This module reads either form from an output directory and prints short text
tables, and turns two counter runs (FETCH_SIZE, then WRITE_SIZE, KiB per
dispatch) plus the same two runs of the plain copy kernel into per-kernel HBM
bytes for bench.py's roofline ``traffic`` field.

    python -m lighthugger_amd.profsum stats DIR [--title T] [--command C]
    python -m lighthugger_amd.profsum counters DIR [--match KERNEL]
    python -m lighthugger_amd.profsum traffic --fetch DIR --write DIR \\
        --copy-fetch DIR --copy-write DIR [--copy-bytes N] [--out FILE]
"""
from __future__ import annotations

import argparse
import csv
import json
import re
import sqlite3
import sys
from pathlib import Path

# Kernel symbols that are unique to one entry of lh_frame_profiled's names.
FRAME_KERNELS = {
    "tile_raster_view": "k_tile_raster<false>",
    "bin_meshlets_view": "k_bin_meshlets<0>",
    "bin_count_view": "k_bin_count",
    "bin_scan_view": "k_bin_scan",
    "bin_fill_view": "k_bin_fill",
    # with the tiled raster on, the per-meshlet kernel of the main view runs as
    # raster_direct_view (the draws the bins left); otherwise as raster_meshlets_view
    "raster_direct_view": "k_raster_meshlets<0, lhd::OneView, false, false>",
    "raster_meshlets_view": "k_raster_meshlets<0, lhd::OneView, false, false>",
    "raster_meshlets_shadow": "k_raster_meshlets<1, lhd::OneView, false, false>",
    "raster_big_view": "k_raster_big<0, lhd::OneView, false, false>",
    "raster_big_shadow": "k_raster_big<1, lhd::OneView, false, false>",
    "raster_clip_view": "k_raster_clip<0, lhd::OneView>",
    "raster_clip_shadow": "k_raster_clip<1, lhd::OneView>",
    "resolve": "k_resolve",
    "depth_bounds": "k_depth_bounds<lhd::OneView>",
    "fit_cull_instances_shadow": "k_fit_cull_instances_shadow<lhd::OneView>",
    "copy": "k_copy",
}


def _short(name: str) -> str:
    """Kernel name without the anonymous namespace and the parameter list."""
    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"^void ", "", n)
    depth, out = 0, []
    for ch in n:  # cut at the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out)


def _files(d: Path, pattern: str) -> list[Path]:
    return sorted(p for p in Path(d).rglob(pattern) if p.is_file())


def kernel_stats(d: Path) -> list[dict]:
    """[{name, calls, total_ns, avg_ns, pct}] sorted by total time."""
    rows = []
    for db in _files(d, "*.db"):
        con = sqlite3.connect(str(db))
        try:
            # the database's top_kernels view reports microseconds
            cur = con.execute("SELECT name, total_calls, total_duration, average, percentage FROM top_kernels")
            rows += [dict(name=r[0], calls=int(r[1]), total_ns=float(r[2]) * 1e3, avg_ns=float(r[3]) * 1e3,
                          pct=float(r[4])) for r in cur]
        except sqlite3.Error:
            pass
        finally:
            con.close()
    for f in _files(d, "*kernel_stats.csv"):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append(dict(name=r["Name"], calls=int(r["Calls"]), total_ns=float(r["TotalDurationNs"]),
                                 avg_ns=float(r["AverageNs"]), pct=float(r["Percentage"])))
    return sorted(rows, key=lambda r: -r["total_ns"])


def counters(d: Path) -> dict[str, dict[str, list[float]]]:
    """{kernel name: {counter: [value per dispatch]}}."""
    out: dict[str, dict[str, list[float]]] = {}
    for db in _files(d, "*.db"):
        con = sqlite3.connect(str(db))
        try:
            cur = con.execute("SELECT kernel_name, counter_name, value FROM counters_collection")
            for k, c, v in cur:
                out.setdefault(k, {}).setdefault(c, []).append(float(v))
        except sqlite3.Error:
            pass
        finally:
            con.close()
    for f in _files(d, "*counter_collection.csv"):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                out.setdefault(r["Kernel_Name"], {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
    return out


def _mean(v: list[float]) -> float:
    return sum(v) / len(v) if v else 0.0


def counter_means(d: Path, counter: str) -> dict[str, tuple[float, int]]:
    """{short kernel name: (mean value per dispatch, dispatches)}."""
    res = {}
    for k, cs in counters(d).items():
        if counter in cs:
            res[_short(k)] = (_mean(cs[counter]), len(cs[counter]))
    return res


def _lookup(table: dict, short: str):
    return table.get(short)


def print_stats(d: Path, title: str, command: str, limit: int = 25) -> str:
    rows = kernel_stats(d)
    lines = [f"# {title}", "", f"command: {command}", "",
             f"{'kernel':<64} {'calls':>7} {'avg_us':>10} {'total_ms':>10} {'pct':>6}"]
    for r in rows[:limit]:
        lines.append(f"{_short(r['name'])[:64]:<64} {r['calls']:>7} {r['avg_ns'] / 1e3:>10.2f} "
                     f"{r['total_ns'] / 1e6:>10.3f} {r['pct']:>6.2f}")
    return "\n".join(lines) + "\n"


def print_counters(d: Path, title: str, command: str, match: str = "") -> str:
    """Mean value per dispatch of every counter collected in `d`, per kernel
    (kernels whose short name contains `match`)."""
    lines = [f"# {title}", "", f"command: {command}", ""]
    for k, cs in sorted(counters(d).items(), key=lambda kv: _short(kv[0])):
        name = _short(k)
        if match and match not in name:
            continue
        vals = "  ".join(f"{c}={_mean(v):.6g}" for c, v in sorted(cs.items()))
        lines.append(f"{name[:72]:<72} n={len(next(iter(cs.values())))}  {vals}")
    return "\n".join(lines) + "\n"


def traffic(fetch: Path, write: Path, copy_fetch: Path, copy_write: Path, copy_bytes: int) -> dict:
    """Per frame kernel: HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE
    (KiB), each scaled by bytes / counter of the plain copy kernel."""
    f, w = counter_means(fetch, "FETCH_SIZE"), counter_means(write, "WRITE_SIZE")
    cf, cw = _lookup(counter_means(copy_fetch, "FETCH_SIZE"), "k_copy"), \
        _lookup(counter_means(copy_write, "WRITE_SIZE"), "k_copy")
    if not cf or not cw:
        raise RuntimeError("copy-kernel counters missing")
    kf, kw = copy_bytes / (cf[0] * 1024.0), copy_bytes / (cw[0] * 1024.0)
    out = {"_calibration": {"copy_bytes": copy_bytes, "fetch_kib": cf[0], "write_kib": cw[0],
                            "fetch_factor": round(kf, 4), "write_factor": round(kw, 4)}}
    for name, sym in FRAME_KERNELS.items():
        a, b = _lookup(f, sym), _lookup(w, sym)
        if a is None or b is None:
            continue
        rb, wb = a[0] * 1024.0 * kf, b[0] * 1024.0 * kw
        out[name] = {"read_bytes": int(rb), "write_bytes": int(wb), "bytes": int(rb + wb),
                     "fetch_kib_raw": round(a[0], 1), "write_kib_raw": round(b[0], 1), "dispatches": a[1]}
    return out


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="rocprofv3 kernel-trace statistics")
    s.add_argument("--command", default="")
    s.add_argument("--limit", type=int, default=25)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--title", default="rocprofv3 counters (mean per dispatch)")
    c.add_argument("--command", default="")
    c.add_argument("--match", default="")
    t = sub.add_parser("traffic")
    for k in ("--fetch", "--write", "--copy-fetch", "--copy-write"):
        t.add_argument(k, required=True)
    t.add_argument("--copy-bytes", type=int, default=256 << 20)
    t.add_argument("--out", default="")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(print_stats(Path(a.dir), a.title, a.command, a.limit))
    elif a.cmd == "counters":
        sys.stdout.write(print_counters(Path(a.dir), a.title, a.command, a.match))
    else:
        res = traffic(Path(a.fetch), Path(a.write), Path(a.copy_fetch), Path(a.copy_write), a.copy_bytes)
        text = json.dumps(res, indent=1) + "\n"
        if a.out:
            Path(a.out).write_text(text)
        sys.stdout.write(text)


if __name__ == "__main__":
    main()
