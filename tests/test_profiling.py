"""The rocprofv3 summary helpers on small synthetic results databases (tables `kernels` and
`pmc_events` with the columns the helpers read): per-instantiation rows combine by dispatch
count, and the counter scale comes from the copy of known size."""
import sqlite3

import pytest

from astaroth_amd import profiling as P


def make_db(path, kernels=(), pmc=()):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, duration real)")
    con.execute("create table pmc_events (name text, counter_name text, counter_value real)")
    con.executemany("insert into kernels values (?, ?)", kernels)
    con.executemany("insert into pmc_events values (?, ?, ?)", pmc)
    con.commit()
    con.close()
    return str(path)


def test_group_stats_combines_instantiations(tmp_path):
    db = make_db(tmp_path / "k.db", kernels=[("solve<float, 0>", 600e3), ("solve<float, 0>", 620e3),
                                              ("solve<float, 1>", 700e3), ("fill", 10e3)])
    r = P.group_stats(db, "solve<float")
    assert r["kernels"] == 2 and r["calls"] == 3
    assert r["avg_us"] == pytest.approx((600 + 620 + 700) / 3)
    assert r["min_us"] == pytest.approx(600) and r["max_us"] == pytest.approx(700)
    with pytest.raises(KeyError):
        P.group_stats(db, "absent")


def test_traffic_weights_by_dispatches_and_scales_by_copy(tmp_path):
    kib = 1024.0
    fetch = make_db(tmp_path / "f.db", pmc=[
        ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 512.0),   # half of the 1 MiB copy counted
        ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 3.0),     # a small copy of the same run
        ("solve<0>", "FETCH_SIZE", 100.0), ("solve<0>", "FETCH_SIZE", 100.0),
        ("solve<2>", "FETCH_SIZE", 400.0)])
    write = make_db(tmp_path / "w.db", pmc=[
        ("__amd_rocclr_copyBuffer", "WRITE_SIZE", 1024.0),
        ("__amd_rocclr_copyBuffer", "WRITE_SIZE", 6.0),
        ("solve<0>", "WRITE_SIZE", 300.0), ("solve<0>", "WRITE_SIZE", 300.0),
        ("solve<2>", "WRITE_SIZE", 0.0)])
    r = P.traffic(fetch, write, int(1024 * kib), "solve<")
    assert r["dispatches"] == 3
    assert r["fetch_scale"] == pytest.approx(2.0) and r["write_scale"] == pytest.approx(1.0)
    assert r["fetch_bytes_per_launch"] == pytest.approx((100 + 100 + 400) / 3 * kib * 2.0)
    assert r["write_bytes_per_launch"] == pytest.approx((300 + 300 + 0) / 3 * kib)
    assert r["bytes_per_launch"] == pytest.approx(r["fetch_bytes_per_launch"] + r["write_bytes_per_launch"])
