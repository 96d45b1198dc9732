"""Summaries of rocprofv3 result databases (rocpd SQLite files) for the benchmark workloads.

    python -m hmpc_amd.profiling kernels RESULTS.db [--symbols a,b] > summary.txt
        per kernel symbol and grid size: dispatches, average / min / max duration

    python -m hmpc_amd.profiling traffic FETCH.db WRITE.db --workload c2 --copy-bytes N [--out traffic.json]
        HBM bytes per dispatch of the workload's kernels from a FETCH_SIZE run and a WRITE_SIZE
        run, each scaled by what the same counter reads for the device copy of N bytes in the
        same run (the copy reads N and writes N bytes)

    python -m hmpc_amd.profiling counters RESULTS.db --workload c2 [--clock-ghz 2.4] [--out valu.json]
        instruction counters (SQ_INSTS_*) per dispatch of the workload's kernels, and the time
        their VALU instructions need at the issue rate of one wave instruction per SIMD every
        4 cycles (CUs x 4 SIMDs x clock / 4); with GRBM_GUI_ACTIVE in the run, also the engine
        clock each kernel ran at (busy cycles per XCD / dispatch duration). Several runs (databases of
        different counters) merge into one entry per kernel in --out.

    python -m hmpc_amd.profiling overlap RESULTS.db [--gap-us 20] [--min-dispatches 20]
        busy periods of a kernel trace (dispatches separated by less than the gap): per period the
        dispatch count, the streams and queues used, the sum of kernel durations, the wall span
        from first start to last end, and their ratio (above 1 where kernels ran concurrently)

The traffic file maps bench kernel names (the names of the library's launch probes) to bytes
per launch; bench.py reports it as roofline.traffic.
"""

from __future__ import annotations

import argparse
import collections
import json
import sqlite3
import sys

# launch-probe name (bench.py) -> kernel symbol prefix of the workload's dispatches
WORKLOAD_KERNELS = {
    "c2": {
        # the streaming column passes (persistent grids: the timed batch is the longest size class)
        "fast polymul forward columns": "hmpc_amd::fast::k_polymul_cols_stream<",
        "fast polymul rows": "hmpc_amd::fast::k_polymul<",        # the two-pass product (N > 4096)
        "fast polymul inverse columns": "hmpc_amd::fast::k_icols_stream<",
    },
    "c3": {
        "fast bgv columns": "hmpc_amd::fast::k_bgv_cols<",
        "fast bgv rows": "hmpc_amd::fast::k_bgv_rows<",
    },
}
COPY_SYMBOL = "__amd_rocclr_copyBuffer"


def _short(name: str) -> str:
    name = name.split("(")[0]
    return name[5:] if name.startswith("void ") else name


def _size_classes(durations):
    """Split one kernel's dispatches of one grid into classes of similar duration (a new class where
    a duration exceeds twice the shortest of the current class): persistent kernels launch the same
    grid for batches of different size (the bench's timed batch and the autotuner's trial batch)."""
    classes = []
    for d in sorted(durations):
        if classes and d <= 2.0 * classes[-1][0]:
            classes[-1].append(d)
        else:
            classes.append([d])
    return classes


def kernel_rows(db_path: str):
    cur = sqlite3.connect(db_path).cursor()
    rows = cur.execute("select name, grid_x, workgroup_x, duration from kernels").fetchall()
    raw = collections.defaultdict(list)
    for name, grid, wg, dur in rows:
        raw[(_short(name), int(grid), int(wg))].append(float(dur))
    groups = {}
    for (name, grid, wg), durs in raw.items():
        classes = _size_classes(durs)
        for i, c in enumerate(classes):
            groups[(name if len(classes) == 1 else f"{name} [size class {i}]", grid, wg)] = c
    return groups


def _longest_class(records):
    """records: (duration, value) per dispatch -> the values of the longest size class"""
    if not records:
        return []
    top = _size_classes([d for d, _ in records])[-1]
    return [v for d, v in records if d >= top[0]]


def counter_rows(db_path: str, counter: str):
    """per (kernel, grid): the counter's value per dispatch of the longest size class"""
    cur = sqlite3.connect(db_path).cursor()
    rows = cur.execute("select kernel_name, grid_size, value, start, end from counters_collection where counter_name = ?",
                       (counter,)).fetchall()
    groups = collections.defaultdict(list)
    for name, grid, value, start, end in rows:
        groups[(_short(name), int(grid))].append((float(end - start), float(value)))
    return {k: _longest_class(v) for k, v in groups.items()}


def cmd_kernels(args):
    groups = kernel_rows(args.db)
    wanted = [s for s in args.symbols.split(",") if s] if args.symbols else None
    print(f"{'kernel':<96} {'grid':>10} {'wg':>5} {'calls':>6} {'avg us':>10} {'min us':>10} {'max us':>10}")
    for (name, grid, wg), d in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        if wanted and not any(name.startswith(w) for w in wanted):
            continue
        print(f"{name[:96]:<96} {grid:>10} {wg:>5} {len(d):>6} {sum(d) / len(d) / 1e3:>10.1f} {min(d) / 1e3:>10.1f} "
              f"{max(d) / 1e3:>10.1f}")


def _largest(groups, prefix):
    cands = [(g, v) for (n, g), v in groups.items() if n.startswith(prefix)]
    if not cands:
        return None, None
    g, v = max(cands, key=lambda c: c[0])
    return g, sum(v) / len(v)


def cmd_traffic(args):
    fetch = counter_rows(args.fetch_db, "FETCH_SIZE")
    write = counter_rows(args.write_db, "WRITE_SIZE")
    _, copy_fetch = _largest(fetch, COPY_SYMBOL)
    _, copy_write = _largest(write, COPY_SYMBOL)
    if not copy_fetch or not copy_write:
        sys.exit("no copy dispatch in the counter runs")
    # counters are KiB per dispatch; the copy moves copy_bytes in each direction
    fscale = args.copy_bytes / (copy_fetch * 1024.0)
    wscale = args.copy_bytes / (copy_write * 1024.0)
    out = {}
    try:
        with open(args.out) as f:
            out = json.load(f)
    except (OSError, ValueError):
        pass
    entries = {}
    lines = [f"copy calibration: {args.copy_bytes} B copy reads FETCH_SIZE {copy_fetch:.0f} KiB, WRITE_SIZE "
             f"{copy_write:.0f} KiB -> scale fetch x{fscale:.3f}, write x{wscale:.3f}",
             f"{'bench kernel':<32} {'grid':>10} {'fetch B':>14} {'write B':>14} {'total B':>14}"]
    for bench_name, prefix in WORKLOAD_KERNELS[args.workload].items():
        g, f = _largest(fetch, prefix)
        _, w = _largest(write, prefix)
        if f is None or w is None:
            continue
        fb, wb = f * 1024.0 * fscale, w * 1024.0 * wscale
        entries[bench_name] = {"bytes_per_launch": fb + wb, "fetch_bytes": fb, "write_bytes": wb, "grid": g,
                               "fetch_scale": fscale, "write_scale": wscale}
        lines.append(f"{bench_name:<32} {g:>10} {fb:>14.0f} {wb:>14.0f} {fb + wb:>14.0f}")
    out[args.workload] = entries
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
            f.write("\n")
    print("\n".join(lines))


def cmd_counters(args):
    cur = sqlite3.connect(args.db).cursor()
    rows = cur.execute("select kernel_name, grid_size, counter_name, value, start, end from counters_collection").fetchall()
    # per (kernel, grid) and counter: (duration, value) per dispatch, then the longest size class
    recs = collections.defaultdict(lambda: collections.defaultdict(list))
    for name, grid, counter, value, start, end in rows:
        recs[(_short(name), int(grid))][counter].append((float(end - start), float(value)))
        if counter == "GRBM_GUI_ACTIVE":
            # GPU busy cycles over the dispatch: with its duration, the engine clock it ran at
            recs[(_short(name), int(grid))]["duration_ns"].append((float(end - start), float(end - start)))
    groups = {k: {c: _longest_class(v) for c, v in cs.items()} for k, cs in recs.items()}
    issue = args.cus * 4 * args.clock_ghz * 1e9 / 4  # wave instructions per second
    entries = {}
    print(f"VALU issue rate: {args.cus} CUs x 4 SIMDs x {args.clock_ghz} GHz / 4 cycles = {issue / 1e9:.1f} G wave-instr/s")
    for bench_name, prefix in WORKLOAD_KERNELS[args.workload].items():
        cands = [(g, c) for (n, g), c in groups.items() if n.startswith(prefix)]
        if not cands:
            continue
        grid, counters = max(cands, key=lambda c: c[0])
        avg = {k: sum(v) / len(v) for k, v in counters.items()}
        print(f"{bench_name} (grid {grid}, {len(next(iter(counters.values())))} dispatches):")
        for k in sorted(avg):
            print(f"  {k:<24} {avg[k]:>16.6g}")
        if "SQ_INSTS_VALU" in avg:
            print(f"  VALU issue time          {avg['SQ_INSTS_VALU'] / issue * 1e3:>13.3f} ms")
        if avg.get("duration_ns"):
            # rocprofv3 sums GRBM_GUI_ACTIVE over the XCDs (8 on MI355X)
            avg["clock_ghz"] = avg["GRBM_GUI_ACTIVE"] / args.xcds / avg["duration_ns"]
            print(f"  engine clock             {avg['clock_ghz']:>13.3f} GHz (GRBM_GUI_ACTIVE / {args.xcds} XCDs / duration)")
        entries[bench_name] = dict(avg, grid=grid)
    if args.out:
        out = {}
        try:
            with open(args.out) as f:
                out = json.load(f)
        except (OSError, ValueError):
            pass
        merged = out.get(args.workload, {})
        for k, v in entries.items():
            merged.setdefault(k, {}).update(v)
        out[args.workload] = merged
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
            f.write("\n")


def cmd_overlap(args):
    cur = sqlite3.connect(args.db).cursor()
    rows = sorted(cur.execute("select start, end, stream_id, queue_id from kernels").fetchall())
    periods, cur_p, last_end = [], [], None
    for start, end, stream, queue in rows:
        if cur_p and start - last_end > args.gap_us * 1e3:
            periods.append(cur_p)
            cur_p = []
        cur_p.append((start, end, stream, queue))
        last_end = end if last_end is None or not cur_p[:-1] else max(last_end, end)
    if cur_p:
        periods.append(cur_p)
    print(f"{'period':>6} {'dispatches':>10} {'streams':>7} {'queues':>6} {'busy ms':>9} {'span ms':>9} {'busy/span':>9}")
    k = 0
    for p in periods:
        if len(p) < args.min_dispatches:
            continue
        busy = sum(e - s for s, e, _, _ in p) / 1e6
        span = (max(e for _, e, _, _ in p) - p[0][0]) / 1e6
        print(f"{k:>6} {len(p):>10} {len({x[2] for x in p}):>7} {len({x[3] for x in p}):>6} {busy:>9.3f} {span:>9.3f} "
              f"{busy / span if span else 0:>9.2f}")
        k += 1


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--symbols", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_db")
    t.add_argument("write_db")
    t.add_argument("--workload", required=True, choices=sorted(WORKLOAD_KERNELS))
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--out", default="")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("--workload", required=True, choices=sorted(WORKLOAD_KERNELS))
    c.add_argument("--cus", type=int, default=256)
    c.add_argument("--clock-ghz", type=float, default=2.4)
    c.add_argument("--xcds", type=int, default=8)
    c.add_argument("--out", default="")
    o = sub.add_parser("overlap")
    o.add_argument("db")
    o.add_argument("--gap-us", type=float, default=20.0)
    o.add_argument("--min-dispatches", type=int, default=20)
    args = ap.parse_args(argv)
    {"kernels": cmd_kernels, "traffic": cmd_traffic, "counters": cmd_counters, "overlap": cmd_overlap}[args.cmd](args)


if __name__ == "__main__":
    main()
