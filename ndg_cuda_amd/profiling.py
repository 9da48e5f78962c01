"""Summaries of rocprofv3 output (rocpd SQLite databases) for the stage kernels.

    python -m ndg_cuda_amd.profiling trace <run_results.db> [--kernel SUBSTR] [--min-us T] [--max-us T] [--grid G]
    python -m ndg_cuda_amd.profiling traffic --fetch DB --write DB --copy-fetch DB --copy-write DB [--grid G] [--last N]
    python -m ndg_cuda_amd.profiling counters DB [DB ...] [--kernel SUBSTR] [--by-grid]

`trace` prints per-kernel call counts and durations from a --kernel-trace run. `traffic` turns one
FETCH_SIZE run and one WRITE_SIZE run of the benchmark into HBM bytes per launch of the kernel,
each scaled by what the same counter reads for a plain 1 GiB device copy (bench.py
--copy-calibration): scale = bytes the copy moves / counter reading for the copy.
"""
import argparse
import glob
import json
import os
import sqlite3
import sys


def _db(path):
    if os.path.isdir(path):
        found = glob.glob(os.path.join(path, "**", "*.db"), recursive=True)
        if not found:
            raise SystemExit(f"no .db under {path}")
        path = found[0]
    return sqlite3.connect(path)


def kernel_stats(path, kernel=None, min_us=0.0, max_us=float("inf"), grid=None):
    """Per-kernel call counts and durations; with `grid`, only dispatches of that grid size
    (threads), i.e. one launch shape."""
    c = _db(path)
    sql = ("select name, count(*), avg(duration), min(duration), max(duration), sum(duration) "
           "from kernels where duration >= ? and duration <= ?")
    args = [min_us * 1e3, min(max_us * 1e3, 1e18)]
    if grid is not None:
        sql += " and grid_x * grid_y * grid_z = ?"
        args.append(int(grid))
    rows = c.execute(sql + " group by name order by sum(duration) desc", args).fetchall()
    out = []
    for name, n, avg, mn, mx, tot in rows:
        if kernel and kernel not in name:
            continue
        out.append({"kernel": name, "calls": n, "avg_ns": avg, "min_ns": mn, "max_ns": mx, "total_ns": tot})
    return out


def counter_means(path, kernel, grid=None, last=None):
    """Mean counter value per dispatch of the kernel; with `grid`, only dispatches of that grid
    size (threads), i.e. one launch shape (the autotuner launches several); with `last`, only the
    last that many dispatches (the timed steps after the tuner's probes, which are mostly
    first-stage launches that read no RK register)."""
    c = _db(path)
    sql = "select counter_name, dispatch_id, value from counters_collection where kernel_name like ?"
    args = [f"%{kernel}%"]
    if grid is not None:
        sql += " and grid_size = ?"
        args.append(int(grid))
    per = {}
    for name, did, v in c.execute(sql + " order by dispatch_id", args).fetchall():
        per.setdefault(name, {}).setdefault(did, 0.0)
        per[name][did] += v
    out = {}
    for name, by in per.items():
        vals = [by[d] for d in sorted(by)]
        if last:
            vals = vals[-last:]
        out[name] = {"dispatches": len(vals), "mean": sum(vals) / len(vals)}
    return out


def grid_sizes(path, kernel):
    """(grid size, dispatches) of every launch shape of the kernel in a counter run."""
    c = _db(path)
    return c.execute("select grid_size, count(distinct dispatch_id) from counters_collection "
                     "where kernel_name like ? group by grid_size order by grid_size", (f"%{kernel}%",)).fetchall()


def traffic(fetch, write, copy_fetch, copy_write, kernel, copy_kernel="copyBuffer", copy_bytes=1 << 30, grid=None,
            last=None):
    f = counter_means(fetch, kernel, grid, last)["FETCH_SIZE"]["mean"]
    w = counter_means(write, kernel, grid, last)["WRITE_SIZE"]["mean"]
    cf = counter_means(copy_fetch, copy_kernel)["FETCH_SIZE"]["mean"]
    cw = counter_means(copy_write, copy_kernel)["WRITE_SIZE"]["mean"]
    fs, ws = copy_bytes / (cf * 1024.0), copy_bytes / (cw * 1024.0)
    read_b, write_b = f * 1024.0 * fs, w * 1024.0 * ws
    return {"kernel": kernel, "fetch_kib_per_dispatch": f, "write_kib_per_dispatch": w,
            "copy_fetch_kib": cf, "copy_write_kib": cw, "fetch_scale": fs, "write_scale": ws,
            "hbm_read_bytes_per_launch": read_b, "hbm_write_bytes_per_launch": write_b,
            "hbm_bytes_per_launch": read_b + write_b}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--kernel")
    t.add_argument("--min-us", type=float, default=0.0, help="only dispatches at least this long")
    t.add_argument("--max-us", type=float, default=float("inf"), help="only dispatches at most this long")
    t.add_argument("--grid", type=int, help="only dispatches of this grid size (threads)")
    tr = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        tr.add_argument("--" + k.replace("_", "-"), required=True)
    tr.add_argument("--kernel", default="stage_kernel<8, false>")
    tr.add_argument("--grid", type=int, help="only dispatches of this grid size (threads)")
    tr.add_argument("--last", type=int, help="only the last this many dispatches (the timed steps)")
    cc = sub.add_parser("counters")
    cc.add_argument("dbs", nargs="+")
    cc.add_argument("--kernel", default="stage_kernel<8, false>")
    cc.add_argument("--by-grid", action="store_true", help="one block per launch shape (grid size)")
    cc.add_argument("--grid", type=int, help="only dispatches of this grid size (threads)")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        for r in kernel_stats(a.db, a.kernel, a.min_us, a.max_us, a.grid):
            print(f"{r['calls']:6d} calls  avg {r['avg_ns'] / 1e3:10.2f} us  min {r['min_ns'] / 1e3:10.2f}  "
                  f"max {r['max_ns'] / 1e3:10.2f}  total {r['total_ns'] / 1e6:9.3f} ms  {r['kernel'][:90]}")
    elif a.cmd == "traffic":
        print(json.dumps(traffic(a.fetch, a.write, a.copy_fetch, a.copy_write, a.kernel, grid=a.grid, last=a.last),
                         indent=1))
    else:
        for db in a.dbs:
            if a.grid is not None:
                grids = [(g, n) for g, n in grid_sizes(db, a.kernel) if g == a.grid]
            elif a.by_grid:
                grids = grid_sizes(db, a.kernel)
            else:
                grids = [(None, None)]
            for g, n in grids:
                if g is not None:
                    print(f"-- grid {g} threads ({n} dispatches)")
                for name, v in sorted(counter_means(db, a.kernel, g).items()):
                    print(f"{name:28s} {v['mean']:.6g}  ({v['dispatches']} dispatches)")
    return 0


if __name__ == "__main__":
    try:
        rc = main()
        sys.stdout.flush()
    except BrokenPipeError:
        # the reader of stdout (head, a pager) closed early: not an error of the summary
        os.dup2(os.open(os.devnull, os.O_WRONLY), sys.stdout.fileno())
        rc = 0
    sys.exit(rc)
