"""The rocprofv3 summary tool (ndg_cuda_amd/profiling.py) behind the commands in profiles/, on
small synthetic databases with the rocpd views it reads (CPU only)."""
import sqlite3

import pytest

from ndg_cuda_amd import profiling


@pytest.fixture
def trace_db(tmp_path):
    path = tmp_path / "trace.db"
    c = sqlite3.connect(path)
    c.execute("create table kernels (name text, duration real, grid_x int, grid_y int, grid_z int)")
    rows = [("stage_kernel<8, false>", 480e3, 7190400, 1, 1)] * 4 + [("stage_kernel<8, false>", 1000e3, 12582912, 1, 1)] \
        + [("mortar_kernel<double, 13>", 8e3, 251648, 1, 1)] * 2
    c.executemany("insert into kernels values (?, ?, ?, ?, ?)", rows)
    c.commit()
    return path


@pytest.fixture
def counter_db(tmp_path):
    def make(name, rows):
        path = tmp_path / name
        c = sqlite3.connect(path)
        c.execute("create table counters_collection (kernel_name text, grid_size int, dispatch_id int, "
                  "counter_name text, value real)")
        c.executemany("insert into counters_collection values (?, ?, ?, ?, ?)", rows)
        c.commit()
        return path
    return make


def test_trace_stats_and_grid_filter(trace_db):
    all_ = profiling.kernel_stats(trace_db, "stage_kernel")
    assert all_[0]["calls"] == 5 and all_[0]["avg_ns"] == pytest.approx((4 * 480e3 + 1000e3) / 5)
    tuned = profiling.kernel_stats(trace_db, "stage_kernel", grid=7190400)
    assert tuned[0]["calls"] == 4 and tuned[0]["avg_ns"] == 480e3
    assert profiling.kernel_stats(trace_db, "stage_kernel", min_us=600)[0]["calls"] == 1
    assert [r["kernel"] for r in profiling.kernel_stats(trace_db)] == ["stage_kernel<8, false>",
                                                                      "mortar_kernel<double, 13>"]


def test_counters_by_grid_and_traffic_calibration(counter_db):
    k = "stage_kernel<8, false>"
    db = counter_db("c.db", [(k, 7190400, 1, "SQ_WAVES", 100.0), (k, 7190400, 2, "SQ_WAVES", 300.0),
                             (k, 12582912, 3, "SQ_WAVES", 1000.0)])
    assert profiling.counter_means(db, k)["SQ_WAVES"]["mean"] == pytest.approx(1400.0 / 3)
    assert profiling.counter_means(db, k, 7190400)["SQ_WAVES"] == {"dispatches": 2, "mean": 200.0}
    assert profiling.grid_sizes(db, k) == [(7190400, 2), (12582912, 1)]
    # a copy of 1 GiB that reads FETCH_SIZE = 512 MiB (scale 2) and writes WRITE_SIZE = 1 GiB (scale 1)
    fetch = counter_db("f.db", [(k, 1, 1, "FETCH_SIZE", 1000.0)])
    write = counter_db("w.db", [(k, 1, 1, "WRITE_SIZE", 3000.0)])
    cfetch = counter_db("cf.db", [("__amd_rocclr_copyBuffer", 1, 1, "FETCH_SIZE", 524288.0)])
    cwrite = counter_db("cw.db", [("__amd_rocclr_copyBuffer", 1, 1, "WRITE_SIZE", 1048576.0)])
    t = profiling.traffic(fetch, write, cfetch, cwrite, k)
    assert t["fetch_kib_per_dispatch"] == 1000.0 and t["write_kib_per_dispatch"] == 3000.0
    assert t["hbm_bytes_per_launch"] == pytest.approx(1000 * 1024 * 2 + 3000 * 1024)


def test_counters_cli_grid_filter_and_closed_pipe(counter_db, capsys):
    import subprocess
    import sys
    k = "stage_kernel<8, false>"
    db = counter_db("g.db", [(k, 7190400, 1, "SQ_WAVES", 100.0), (k, 12582912, 2, "SQ_WAVES", 1000.0)])
    assert profiling.main(["counters", str(db), "--kernel", k, "--grid", "12582912"]) == 0
    out = capsys.readouterr().out
    assert "-- grid 12582912 threads (1 dispatches)" in out and "1000" in out and "7190400" not in out
    # a reader that closes the pipe after one line does not turn the summary into a failure
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    many = " ".join([db.name] * 2000)  # about 300 KB of summary, more than a pipe holds
    env = dict(__import__("os").environ, PYTHONPATH=root)
    # pipefail: the exit status is Python's (head's would always be 0)
    r = subprocess.run(f"set -o pipefail; {sys.executable} -m ndg_cuda_amd.profiling counters {many} --kernel '{k}'"
                       " --by-grid | head -1", shell=True, cwd=db.parent, env=env, capture_output=True, text=True,
                       executable="/bin/bash")
    assert r.returncode == 0 and "Traceback" not in r.stderr
