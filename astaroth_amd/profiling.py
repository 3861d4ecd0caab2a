"""Summaries of rocprofv3 results databases (SQLite, rocprofv3 >= ROCm 7) for the MHD hot path.

    python -m astaroth_amd.profiling stats  <results.db>            # kernel time summary
    python -m astaroth_amd.profiling group  <results.db> <kernel-substring>  # one line for all matches
    python -m astaroth_amd.profiling traffic <fetch.db> <write.db> <copy_bytes> <kernel-substring>

A substring can match several instantiations of one kernel (the FP32 solve has one per RK3
substep); `group` and `traffic` then combine them, weighted by their dispatch counts.

`traffic` reads FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the matching kernels, scales each
by what the same counter reads for the device-to-device copy of `copy_bytes` bytes recorded in
the same run (a kernel whose true traffic is copy_bytes read + copy_bytes written; the largest
copy dispatch of the run), and prints a
JSON record with HBM bytes per launch.
"""
import json
import sqlite3
import sys


def kernel_stats(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                       "max(duration) from kernels group by name order by sum(duration) desc")
    out = []
    for name, n, tot, avg, mn, mx in rows:
        out.append(dict(name=name, calls=n, total_us=tot / 1e3, avg_us=avg / 1e3,
                        min_us=mn / 1e3, max_us=mx / 1e3))
    return out


def counter_means(db, counter):
    con = sqlite3.connect(db)
    rows = con.execute("select name, count(*), avg(counter_value) from pmc_events "
                       "where counter_name = ? group by name", (counter,))
    return {name: (n, val) for name, n, val in rows}


def counter_max(db, counter, sub):
    """The largest per-dispatch value of a counter over the kernels whose name contains sub."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, counter_value from pmc_events where counter_name = ?",
                       (counter,))
    vals = [v for name, v in rows if sub in name]
    if not vals:
        raise KeyError(f"no kernel matching {sub!r}")
    return max(vals)


def traffic(fetch_db, write_db, copy_bytes, kernel_sub):
    f = counter_means(fetch_db, "FETCH_SIZE")
    w = counter_means(write_db, "WRITE_SIZE")

    def pick(d, sub):
        """(matching names, dispatches, dispatch-weighted mean) over every kernel matching sub"""
        keys = sorted(k for k in d if sub in k)
        if not keys:
            raise KeyError(f"no kernel matching {sub!r}: {sorted(d)}")
        n = sum(d[k][0] for k in keys)
        return keys, n, sum(d[k][0] * d[k][1] for k in keys) / n

    # the calibration copy is the largest copy of the run (a run may hold many small ones)
    copy_f = counter_max(fetch_db, "FETCH_SIZE", "copyBuffer")
    copy_w = counter_max(write_db, "WRITE_SIZE", "copyBuffer")
    kf, nf, vf = pick(f, kernel_sub)
    _, _, vw = pick(w, kernel_sub)
    scale_f = copy_bytes / (copy_f * 1024.0)
    scale_w = copy_bytes / (copy_w * 1024.0)
    fetch_b = vf * 1024.0 * scale_f
    write_b = vw * 1024.0 * scale_w
    return dict(kernel=kf if len(kf) > 1 else kf[0], dispatches=nf, fetch_kib_raw=vf,
                write_kib_raw=vw, copy_bytes=copy_bytes, copy_fetch_kib=copy_f,
                copy_write_kib=copy_w, fetch_scale=scale_f, write_scale=scale_w,
                fetch_bytes_per_launch=fetch_b, write_bytes_per_launch=write_b,
                bytes_per_launch=fetch_b + write_b)


def group_stats(db, sub):
    """One summary over every kernel whose name contains sub (calls, total, mean, min, max)."""
    rows = [r for r in kernel_stats(db) if sub in r["name"]]
    if not rows:
        raise KeyError(f"no kernel matching {sub!r}")
    n = sum(r["calls"] for r in rows)
    tot = sum(r["total_us"] for r in rows)
    return dict(name=sub, kernels=len(rows), calls=n, total_us=tot, avg_us=tot / n,
                min_us=min(r["min_us"] for r in rows), max_us=max(r["max_us"] for r in rows))


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        for r in kernel_stats(argv[1]):
            print(f"| {r['name'][:90]} | {r['calls']} | {r['total_us']:.1f} | {r['avg_us']:.1f} | "
                  f"{r['min_us']:.1f} | {r['max_us']:.1f} |")
        return 0
    if len(argv) >= 3 and argv[0] == "group":
        r = group_stats(argv[1], argv[2])
        print(f"| {r['name']} ({r['kernels']} kernels) | {r['calls']} | {r['total_us']:.1f} | "
              f"{r['avg_us']:.1f} | {r['min_us']:.1f} | {r['max_us']:.1f} |")
        return 0
    if len(argv) >= 5 and argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], int(argv[3]), argv[4]), indent=1))
        return 0
    print(__doc__)
    return 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
