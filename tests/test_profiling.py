"""The rocprofv3 summaries behind profiles/ (ptrans.profiling), on small SQLite files laid
out like rocprofv3's result databases: per-kernel durations, the span of a launch made of
concurrent kernels, and per-launch counter sums scaled by a device copy's counters."""
import sqlite3

import pytest

from ptrans import profiling


def _kt(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, start integer, end integer)")
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.commit()
    con.close()
    return str(path)


def _pmc(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, dispatch_id integer, "
                "counter_name text, value real, start integer, end integer)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?, 0, 0)", rows)
    con.commit()
    con.close()
    return str(path)


def test_kernel_stats_average_and_share(tmp_path):
    db = _kt(tmp_path / "kt.db", [("k1", 0, 100), ("k1", 200, 500), ("k2", 600, 700)])
    st = {k["name"]: k for k in profiling.kernel_stats(db)}
    assert st["k1"]["calls"] == 2 and st["k1"]["average_ns"] == 200 and st["k1"]["total_ns"] == 400
    assert st["k2"]["percent"] == pytest.approx(20.0)


def test_launch_span_of_concurrent_kernels(tmp_path):
    # two launches, each: a side kernel started just before the marker kernel and running
    # beside it, then a kernel after it; a foreign kernel between them is not counted
    ms = 1_000_000
    rows = [("side", 0, 50 * ms), ("mark", 1 * ms, 30 * ms), ("after", 31 * ms, 60 * ms),
            ("other", 61 * ms, 90 * ms),
            ("side", 99 * ms, 120 * ms), ("mark", 100 * ms, 140 * ms), ("after", 141 * ms, 150 * ms)]
    db = _kt(tmp_path / "kt.db", rows)
    avg, n = profiling.launch_spans(db, ["mark", "side", "after"], lead_ns=2 * ms)
    assert n == 2
    assert avg == pytest.approx((60 * ms + 51 * ms) / 2)
    # a sum of the kernels' durations would count the overlap twice
    assert avg < (50 + 29 + 29 + 21 + 40 + 9) * ms / 2


def test_traffic_summary_sums_counters_per_launch(tmp_path):
    ms = 1_000_000
    kt = _kt(tmp_path / "kt.db", [("side", 0, 50 * ms), ("mark", 1 * ms, 30 * ms),
                                  ("side", 31 * ms, 40 * ms)])
    # per dispatch, in KiB: the launch has one "mark" dispatch and two "side" dispatches
    f = _pmc(tmp_path / "f.db", [("mark", 1, "FETCH_SIZE", 10.0), ("side", 2, "FETCH_SIZE", 3.0),
                                 ("side", 3, "FETCH_SIZE", 5.0)])
    w = _pmc(tmp_path / "w.db", [("mark", 1, "WRITE_SIZE", 20.0), ("side", 2, "WRITE_SIZE", 1.0),
                                 ("side", 3, "WRITE_SIZE", 1.0)])
    sq = _pmc(tmp_path / "sq.db", [("mark", 1, "SQ_INSTS_VALU", 1000.0),
                                   ("side", 2, "SQ_INSTS_VALU", 10.0),
                                   ("side", 3, "SQ_INSTS_VALU", 30.0)])
    # a 1 MiB copy reads 512 KiB of FETCH_SIZE (scale 2) and 1024 KiB of WRITE_SIZE (scale 1)
    cf = _pmc(tmp_path / "cf.db", [("__amd_rocclr_copyBuffer", 1, "FETCH_SIZE", 512.0)])
    cw = _pmc(tmp_path / "cw.db", [("__amd_rocclr_copyBuffer", 1, "WRITE_SIZE", 1024.0)])
    out = profiling.traffic_summary(kt, f, w, cf, cw, sq, 1.0, 100, "toy",
                                    {"launch": "mark+side", "alone": "mark"})
    k = out["kernels"]["launch"]
    assert out["fetch_scale"] == pytest.approx(2.0) and out["write_scale"] == pytest.approx(1.0)
    assert k["average_us"] == pytest.approx(50 * ms / 1e3)
    assert k["hbm_fetch_bytes"] == pytest.approx((10 + 3 + 5) * 1024 * 2.0)
    assert k["hbm_write_bytes"] == pytest.approx((20 + 1 + 1) * 1024)
    assert k["valu_wave_instr_per_path"] == pytest.approx(1040 / 100)
    a = out["kernels"]["alone"]
    assert a["average_us"] == pytest.approx(29 * ms / 1e3)
    assert a["hbm_fetch_bytes"] == pytest.approx(10 * 1024 * 2.0)
